// panslbm2_amd/csrc/bc3d.h -- per-node boundary operators for the 3D lattices (device).
//
// Face numbering: 0 xmin, 1 xmax, 2 ymin, 3 ymax, 4 zmin, 5 zmax.  The host API rejects 3D
// boundary operators that are not implemented here yet, so these bodies are never reached with
// other operators.
#pragma once
#include "lattice.h"

namespace lbm {

__device__ __forceinline__ void ns_set_u_3d(double *, int, double, double, double) {}
__device__ __forceinline__ void ns_set_rho_3d(double *, int, double, double, double) {}
template <class L> __device__ __forceinline__ void ad_set_t_3d(double *, int, double, double, double, double) {}
template <class L> __device__ __forceinline__ void ad_set_q_3d(double *, int, double, double, double, double, double) {}
__device__ __forceinline__ void ans_iset_u_3d(double *, int, double, double, double, double) {}
__device__ __forceinline__ void ans_iset_rho_3d(double *, int) {}
template <class L> __device__ __forceinline__ void aad_iset_t_3d(double *, int, double, double, double) {}
template <class L> __device__ __forceinline__ void aad_iset_q_3d(double *, int, double, double, double, double) {}
template <class LQ>
__device__ __forceinline__ void aad_iset_rho_3d(double *, const double *, int, int, double, double, double, double,
                                                double, double) {}

}  // namespace lbm
