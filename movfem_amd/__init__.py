"""MI355X-native element integration, assembly and Krylov solve for MoVFEM_3DMT."""
