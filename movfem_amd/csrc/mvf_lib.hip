// Single translation unit for libmovfem_gpu.so.
#include "mvf_setup.hip"
#include "mvf_integrate.hip"
#include "mvf_boundary.hip"
#include "mvf_solver.hip"
#include "mvf_direct.hip"
#include "mvf_readout.hip"
#include "mvf_api.hip"
#include "mvf_output.hip"
