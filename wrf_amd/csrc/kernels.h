// Kernel entry points of the dynamics step, shared by the kernel sources and the host driver.
#pragma once
#include "dev_common.h"

namespace wrfmi {
constexpr int kAdvUvK = 8;  // levels per thread in k_advance_uv
constexpr int kAdvK = 8;     // levels per block in k_adv_scalar
constexpr int kTendK = 8;    // levels per thread in k_tend_u/v/w/t
constexpr int kAdvTileI = 63;  // i points per wave in k_adv_scalar (lane 63 supplies a face)
constexpr int kAdvRows = 4;    // j rows (waves) per block in k_adv_scalar
__global__ void k_step_prep(KArgs);
template <bool COMP> __global__ void k_ww_scan(KArgs);
template <bool PER> __global__ void k_tend_u(KArgs);
template <bool PER> __global__ void k_tend_v(KArgs);
template <bool PER> __global__ void k_tend_w(KArgs);
template <bool PER> __global__ void k_tend_t(KArgs);
__global__ void k_small_prep(KArgs);
__global__ void k_p_rho0(KArgs);
template <typename CT> __global__ void k_coef_w(KArgs);
template <bool PER> __global__ void k_advance_uv(KArgs);
__global__ void k_ac_div(KArgs);
template <bool COMP> __global__ void k_ac_mu_ww(KArgs);
__global__ void k_ac_theta(KArgs);
__global__ void k_ac_wpred(KArgs);
template <typename CT> __global__ void k_ac_wsolve(KArgs);
__global__ void k_small_finish(KArgs);
template <bool PER> __global__ void k_adv_scalar(KArgs);
__global__ void k_update_scalar(KArgs);
__global__ void k_p_rho_phi(KArgs);
__global__ void k_halo(HaloArgs);
__global__ void k_count_nonfinite(const float *a, Dm d, int kend, int is3d, unsigned long long *cnt);
}  // namespace wrfmi
