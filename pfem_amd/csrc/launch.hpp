// Host-side declarations of the kernel launchers (one translation unit per kernel family).
#pragma once

#include <hip/hip_runtime.h>
#include "common.hpp"
#include "views.hpp"

namespace pfem {

void launch_pspg_assemble(int dim, const PspgView* views, const BlockTask* tasks, int ntasks, hipStream_t s);
void launch_pspg_bc(const PspgView* views, const BlockTask* tasks, int ntasks, hipStream_t s);

// ids: the cases (case_id of each entry) whose KScal::req_init is set, nids of them
template <typename T>
void launch_kry_init(int lanes, const KrylovView<T>* v, const BlockTask* t, int nt, const BlockTask* ids, int nids,
                     int par, hipStream_t s);
template <typename T>
void launch_kry_phase(int phase, int lanes, const KrylovView<T>* v, const BlockTask* t, int nt, int par,
                      hipStream_t s, hipEvent_t* ev = nullptr);
template <typename T>
void launch_spmv(int lanes, const NodeMat<T>& A, const T* x, T* y, int nblocks, hipStream_t s);
// y = A x over the sliced ELL copy
void launch_spmv_ell(const EllMatT<double>& E, const double* x, double* y, int nblocks, hipStream_t s);

// aggregation multigrid (T: FP64 or FP32 hierarchy)
template <typename T>
void launch_amg_galerkin(const AmgView<T>* v, const BlockTask* t, int nt, hipStream_t s);
template <typename T>
void launch_amg_bjac(const AmgView<T>* v, const BlockTask* t, int nt, hipStream_t s);
template <typename T>
void launch_amg_dense(const AmgView<T>* v, const BlockTask* t, int nt, hipStream_t s);
template <typename T>
void launch_amg_pre(const AmgView<T>* v, const BlockTask* t, int nt, int which, hipStream_t s);
template <typename T>
void launch_amg_res(const AmgView<T>* v, const BlockTask* t, int nt, int which, hipStream_t s);
template <typename T>
void launch_amg_restrict(const AmgView<T>* v, const BlockTask* t, int nt, hipStream_t s);
template <typename T>
void launch_amg_coarsest(const AmgView<T>* v, const BlockTask* t, int nt, hipStream_t s);
template <typename T>
void launch_amg_ap(const AmgView<T>* v, const BlockTask* t, int nt, hipStream_t s);
template <typename T>
void launch_amg_post(const AmgView<T>* v, const BlockTask* t, int nt, int which, hipStream_t s);
void launch_ir_residual(int lanes, const RefineView* v, const BlockTask* t, int nt, hipStream_t s);
void launch_ir_update(const RefineView* v, const BlockTask* t, int nt, hipStream_t s);

void launch_pic_copy(const PicardView* v, const BlockTask* t, int nt, hipStream_t s);
void launch_update_pos(const PicardView* v, const BlockTask* t, int nt, hipStream_t s);
void launch_res_axf(int lanes, const PicardView* v, const BlockTask* t, int nt, hipStream_t s);
void launch_res_u(const PicardView* v, const BlockTask* t, int nt, hipStream_t s);

// node_tasks: one lane per node; gather2 / gather3: kWcLanes lanes per node of 2D / 3D cases;
// elem2 / elem3: one lane per element of 2D / 3D cases
template <typename S>
void launch_wc_step(const WcView<S>* v, const BlockTask* node_tasks, int n_node_tasks, const BlockTask* gather2,
                    int n_gather2, const BlockTask* gather3, int n_gather3, const BlockTask* elem2, int n_elem2,
                    const BlockTask* elem3, int n_elem3, hipStream_t s);
template <typename S>
void launch_wc_cfl(const WcView<S>* v, const BlockTask* elem2, int n_elem2, const BlockTask* elem3, int n_elem3,
                   hipStream_t s);

// Extractors: one partial per workgroup (nblocks of them) into partials.  rho: null (PSPG)
// or the density array in FP64 / FP32 (rho_fp32).
void launch_mass(int dim, int N, int E, const int* conn, const double* pos, const void* rho, int rho_fp32,
                 double* partials, int nblocks, hipStream_t s);
void launch_minmax(int N, const uint8_t* flags, const double* x, int want_max, double* partials, int nblocks,
                   hipStream_t s);
// out[0] = partials[0] + ... + partials[n-1], added in that order (one thread).
void launch_sum_partials(const double* partials, int n, double* out, hipStream_t s);

// The fixed device scratch of pfem_case_mass_noalloc: kScratchPartials doubles per GPU.
// init_scratch resolves its address on `device` (which must be current), once per device;
// scratch_partials returns that address.
constexpr int kScratchPartials = 2048;
constexpr int kMaxDevices = 64;
void init_scratch(int device);
double* scratch_partials(int device);

void launch_stream_read(const double* a, long long n, double* out, int nblocks, hipStream_t s);
void launch_copy(const double* a, double* b, long long n, int nblocks, hipStream_t s);
void launch_xcd_probe(int* out, int nblocks, hipStream_t s);

}  // namespace pfem
