// Augmented-space recursion and KPM moment engine for MI355X (gfx950): HIP kernels and the C ABI of include/asr_hip.h.
//
// Chains that share one node set (one sublattice operator) are stacked as the members of a block vector,
//     V[(node * 9 + orbital) * nbp + member]        (member fastest, as asr_augblk's x(j, 9(i-1)+l), src/asr_augblk.f90:1-13)
// so one pass over the node tables, the bond blocks and the gathered amplitudes serves every member.  A wavefront covers
// npw = 64 / MT nodes x MT members (MT = member tile, a power of two <= 64): neighbour indices and bond blocks are
// wave-uniform per node, and the gathers of one neighbour are contiguous across the lanes of its members.
//
// One chain level of the edc route (src/asr_edc.f90:95-155 / src/asr_augblk.f90:219-250) is
//   K1  h1 = rand_c(u~) + D S D u~                         (apply_h, src/asr_augcls.f90:449-489, D uniform: diagonal)
//   K2  r  = rand_o(h1)                                    (apply_rand, :380-403)
//   K3  w  = sq (rand_e(u~) + h1 - rand_c(r) - D S D r + z rand_pa(u~) + rand_pb(u~)),  partial a = sum u w
//   fin a;  K4  w -= a u + b up, project(w) (augcls_project_c, :329-348), partial sum w w;  fin b;  K5  u = w / b
// with u~ = u / sq (normalised -> amplitude coordinates).  A species parameter that differs between species (d in the plain
// route) takes the general path: t1 = rand_d(u~); t2 = S t1; h = rand_c(u~) + rand_d(t2).  All reductions are two-stage
// (per-block partials in a fixed tree order, then a fixed-order sum), so results are bitwise reproducible run to run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "asr_hip.h"
#include "asr_host.h"

// Device scratch of asr_gpu_kpm_moments_noalloc (64 MiB per device, held from the load of the library to the end of the
// process): the entry point that allocates nothing carves its work arrays from it.  In host code `scratch` names a
// host-side stand-in, not device memory: the device address comes from hipGetSymbolAddress, per device (scratch_ptr).
#define ASR_NOALLOC_SCRATCH_DOUBLES ((size_t)8 << 20)
static __device__ __align__(256) double scratch[ASR_NOALLOC_SCRATCH_DOUBLES];

namespace {

thread_local std::string g_err;

int fail(const std::string& msg) {
  g_err = msg;
  return 1;
}

#define HCHECK(call)                                                                                   \
  do {                                                                                                 \
    hipError_t e_ = (call);                                                                            \
    if (e_ != hipSuccess) return fail(std::string(#call) + ": " + hipGetErrorString(e_));              \
  } while (0)

// The device address of `scratch` on a device (looked up once per device, with that device current; every call on the
// device shares the array, so calls that overlap on one GPU race).
static std::mutex g_scratch_mu;
static double* g_scratch_addr[64] = {};
static int scratch_ptr(int device, double** out) {
  if (device < 0 || device >= 64) return fail("device scratch: device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_addr[device]) {
    int cur = 0;
    HCHECK(hipGetDevice(&cur));
    if (cur != device) HCHECK(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) HCHECK(hipSetDevice(cur));
    if (e != hipSuccess || !p) return fail(std::string("device scratch: hipGetSymbolAddress: ") + hipGetErrorString(e));
    g_scratch_addr[device] = (double*)p;
  }
  *out = g_scratch_addr[device];
  return 0;
}


// parameter slots: c, d, o, e; then pairs (a, b) of energy-linear diagonals X(z) = z X(a) + X(b): pa/pb (route edc's X(z);
// route ed3's z - et), ua/ub (ed3's u), qa/qb (ed3's Qt)
constexpr int NQ = 10;
enum { QC = 0, QD = 1, QO = 2, QE = 3, QPA = 4, QPB = 5, QUA = 6, QUB = 7, QQA = 8, QQB = 9 };
constexpr int WAVE = 64;
constexpr int BLOCK = 256;

// ------------------------------------------------------------------------------------------------ scalar types
struct cplx {
  double re, im;
};
__host__ __device__ inline cplx operator+(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }
__host__ __device__ inline cplx operator-(cplx a, cplx b) { return {a.re - b.re, a.im - b.im}; }
__host__ __device__ inline cplx operator*(double s, cplx a) { return {s * a.re, s * a.im}; }
__host__ __device__ inline cplx operator*(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__host__ __device__ inline cplx operator/(cplx a, double s) { return {a.re / s, a.im / s}; }

template <class T> __device__ inline T zero();
template <> __device__ inline double zero<double>() { return 0.0; }
template <> __device__ inline cplx zero<cplx>() { return {0.0, 0.0}; }
__device__ inline double fma_s(double s, double x, double acc) { return fma(s, x, acc); }
__device__ inline cplx fma_s(double s, cplx x, cplx acc) { return {fma(s, x.re, acc.re), fma(s, x.im, acc.im)}; }
__device__ inline double bilin(double a, double b) { return a * b; }
__device__ inline cplx bilin(cplx a, cplx b) { return a * b; }
__device__ inline double shfl_xor(double v, int m) { return __shfl_xor(v, m); }
__device__ inline cplx shfl_xor(cplx v, int m) { return {__shfl_xor(v.re, m), __shfl_xor(v.im, m)}; }
__device__ inline double inv(double v) { return 1.0 / v; }
__device__ inline cplx inv(cplx v) {
  double d = v.re * v.re + v.im * v.im;
  return {v.re / d, -v.im / d};
}

// ------------------------------------------------------------------------------------------------ device operator
struct OpDev {
  int nn, nslot, nop, nsmax, ncls, nsp, ntype, nstab;
  int piid;              // every operation maps each orbital onto itself (pi = identity)
  const uint32_t* hop;   // nn x nslot: (node + 1) << 6 | op, 0 = no link
  const uint32_t* flip;  // nn x nsmax, same packing
  const uint32_t* info;  // nn: lst | ncl << 4 | ntyp << 12 | fz << 20 (fz: frozen-site slot 0 none, 1 start, 2 partner)
                         //     | INFO_PARTNERS (the node has a flip partner other than itself)
  const double* sq;      // nn + 1 (sq[nn] = 1: the padding node)
  const double* isq;     // nn + 1: 1 / sq
  const uint32_t* sgm;   // nop: bit l set when sg(l, g) < 0
  const double* blkp;    // ntype x (nslot + 1) x 9 x 10: the bond blocks with rows padded to 10 (16-byte aligned rows),
                         // followed by the same blocks transposed (columns padded to 10)
  const int* stab_ptr;   // nn + 1
  const int* stab;
  const int* pi;         // nop x 9
  const double* sg;      // nop x 9
  const double* blk;     // ntype x (nslot + 1) x 81: slot 0 = on-site block
  const int* nsc;        // ncls
};

// per-member tables, member fastest (stride nbp)
struct MemDev {
  int nb, nbp, mt;            // members, padded stride, member tile
  const uint64_t* chimask;    // nbp: bit g set <=> chi(g) = -1
  const double* wts;          // NQ x ncls x nsmax x nsmax x 9 x nbp
  const double* fpar;         // 2 x NQ x 9 x nbp: parameters of the frozen species of slot 1 (start), slot 2 (partner)
  const double* dvec;         // 9 x nbp: d per orbital (uniform-d members)
  const cplx* z;              // nbp
  const double* dperm;        // nop x 9 x nbp: d at the permuted orbital, d[pi(l, g)]
  int pair;                   // complex members as lane pairs (member tile 64 only)
  const int* tact;            // per member tile: 0 = every member has left (the tile's blocks return at once), or null
  const double* mtm;          // ncls x nsmax x nsmax x nsmax x nbp: class tables per member
  const int* fsp;             // 2 x nbp: species of the frozen start site and of the frozen partner site (-1: none)
  int ntl;                    // member tiles launched (the leading ones; 0 = all): after a compaction the running tiles
};

// The frozen-site slot of node info inf for member m: 1 (the start site), 2 (a frozen pair partner, augcls_freeze_any)
// when the member freezes that site, else 0.
constexpr uint32_t INFO_PARTNERS = 1u << 24;

__device__ inline int frozen_slot(uint32_t inf, const MemDev& md, int m) {
  const int s = (int)((inf >> 20) & 3u);
  return (s && md.fsp[(size_t)(s - 1) * md.nbp + m] >= 0) ? s : 0;
}
__device__ inline const double* frozen_row(const MemDev& md, int slot, int q) {
  return md.fpar + ((size_t)(slot - 1) * NQ + q) * 9 * md.nbp;
}

__device__ inline void unpack_link(uint32_t e, int& r, int& g) {
  r = (int)(e >> 6) - 1;
  g = (int)(e & 63u);
}

// Workgroups are dealt round-robin over the 8 XCDs (each with its own L2): map the hardware block index so that every XCD
// processes one contiguous range of logical blocks, which keeps neighbouring sites / nodes in the same L2.
__device__ inline int xcd_block(int bid, int nblk) {
  const int nx = 8, per = nblk / nx, rem = nblk % nx, x = bid % nx, idx = bid / nx;
  return x < rem ? x * (per + 1) + idx : rem * (per + 1) + (x - rem) * per + idx;
}

template <class T>
__device__ inline T ld(const T* x, int node, int l, int nbp, int m) {
  return x[((size_t)node * 9 + l) * nbp + m];
}

// Copies the padded bond blocks of every site type into LDS (all threads of the block), then synchronises.
__device__ __forceinline__ void stage_blocks(const OpDev& op, double* lds) {
  const int n = op.ntype * (op.nslot + 1) * 90;
  for (int q = threadIdx.x; q < n; q += blockDim.x) lds[q] = op.blkp[q];
  __syncthreads();
}

// acc += coef * [dout o] S ([din o] x~)  with x~ = x / sq (xnorm) and the member's characters; B in LDS
template <class T>
__device__ void s_accum(const OpDev& op, const double* __restrict__ blk, const int* __restrict__ pi,
                        const double* __restrict__ sg, uint64_t chim, int i, int t, int m, int nbp, const T* __restrict__ x,
                        int xnorm, const double* dv, const double* __restrict__ dperm, T acc[9]) {
  T v[9];
  const double* b0 = blk + (size_t)t * (op.nslot + 1) * 81;
  {
    const double inv = xnorm ? op.isq[i] : 1.0;
    for (int l = 0; l < 9; ++l) {
      T xv = ld(x, i, l, nbp, m);
      double s = dv ? dv[l] : 1.0;
      v[l] = (s * inv) * xv;
    }
    for (int l = 0; l < 9; ++l) {
      T a = acc[l];
      for (int l2 = 0; l2 < 9; ++l2) a = fma_s(b0[l * 9 + l2], v[l2], a);
      acc[l] = a;
    }
  }
  const uint32_t* hp = op.hop + (size_t)i * op.nslot;
  for (int k = 0; k < op.nslot; ++k) {
    uint32_t e = hp[k];
    if (e == 0u) continue;
    int r, g;
    unpack_link(e, r, g);
    double chi = ((chim >> g) & 1ull) ? -1.0 : 1.0;
    const int* pg = pi + g * 9;
    const double* sgg = sg + g * 9;
    const double isr = xnorm ? op.isq[r] : 1.0;
    for (int l = 0; l < 9; ++l) {
      int p = pg[l];
      T xv = ld(x, r, p, nbp, m);
      // d of the gathered orbital pi(l): from the member's permuted table (a run-time index into dv would spill)
      double s = chi * sgg[l] * (dv ? dperm[((size_t)g * 9 + l) * nbp + m] : 1.0);
      v[l] = (s * isr) * xv;
    }
    const double* bk = b0 + (size_t)(k + 1) * 81;
    for (int l = 0; l < 9; ++l) {
      T a = acc[l];
      for (int l2 = 0; l2 < 9; ++l2) a = fma_s(bk[l * 9 + l2], v[l2], a);
      acc[l] = a;
    }
  }
}

// acc += X(par q) x~ at node i (apply_rand: own state plus flip partners; frozen nodes carry the frozen species)
template <class T>
__device__ void rand_accum(const OpDev& op, const MemDev& md, const int* __restrict__ pi, const double* __restrict__ sg,
                           uint64_t chim, int i, uint32_t inf, int m, int q, const T* __restrict__ x, int xnorm, double coef,
                           T acc[9]) {
  const int nbp = md.nbp, nsm = op.nsmax;
  const int fs = frozen_slot(inf, md, m);
  const double isi = op.isq[i];
  if (fs) {
    const double* fp = frozen_row(md, fs, q);
    for (int l = 0; l < 9; ++l) {
      T xv = ld(x, i, l, nbp, m);
      if (xnorm) xv = isi * xv;
      acc[l] = fma_s(coef * fp[(size_t)l * nbp + m], xv, acc[l]);
    }
    return;
  }
  const int l0 = inf & 15u, kap = (inf >> 4) & 255u;
  const double* w = md.wts + (((size_t)q * op.ncls + kap) * nsm + l0) * nsm * 9 * nbp;
  T r_acc[9];
  for (int l = 0; l < 9; ++l) {
    T xv = ld(x, i, l, nbp, m);
    if (xnorm) xv = isi * xv;
    r_acc[l] = fma_s(w[((size_t)l0 * 9 + l) * nbp + m], xv, zero<T>());
  }
  const int nsc = op.nsc[kap];
  const uint32_t* fl = op.flip + (size_t)i * nsm;
  for (int k = 0; k < nsc; ++k) {
    uint32_t e = fl[k];
    if (k == l0 || e == 0u) continue;
    int r, g;
    unpack_link(e, r, g);
    double chi = ((chim >> g) & 1ull) ? -1.0 : 1.0;
    const double isr = op.isq[r];
    for (int l = 0; l < 9; ++l) {
      T xv = ld(x, r, pi[g * 9 + l], nbp, m);
      if (xnorm) xv = isr * xv;
      T tf = (chi * sg[g * 9 + l]) * xv;
      r_acc[l] = fma_s(w[((size_t)k * 9 + l) * nbp + m], tf, r_acc[l]);
    }
  }
  for (int l = 0; l < 9; ++l) acc[l] = fma_s(coef, r_acc[l], acc[l]);
}

// X(z) x~ = z (X(pa) x~) + X(pb) x~ for complex members (blk_rand_lin / rand_lin_amp, src/asr_augblk.f90:127-165)
__device__ void lin_accum(const OpDev& op, const MemDev& md, const int* pi, const double* sg, uint64_t chim, int i,
                          uint32_t inf, int m, int qa, const cplx* x, int xnorm, double coef, cplx acc[9]) {
  cplx ta[9], tb[9];
  for (int l = 0; l < 9; ++l) { ta[l] = {0.0, 0.0}; tb[l] = {0.0, 0.0}; }
  rand_accum<cplx>(op, md, pi, sg, chim, i, inf, m, qa, x, xnorm, 1.0, ta);
  rand_accum<cplx>(op, md, pi, sg, chim, i, inf, m, qa + 1, x, xnorm, 1.0, tb);
  cplx z = md.z[m];
  for (int l = 0; l < 9; ++l) acc[l] = acc[l] + coef * (z * ta[l] + tb[l]);
}
__device__ inline void lin_accum(const OpDev&, const MemDev&, const int*, const double*, uint64_t, int, uint32_t, int, int,
                                 const double*, int, double, double*) {}

struct FArgs {
  const void* xs;  // S part input (null: no S part)
  int xs_norm, s_din, s_dout;
  double s_coef;
  const void* xr[3];  // rand terms
  int qr[3], r_norm[3];
  double r_coef[3];
  int nr;
  const void* xl[2];  // complex energy-linear terms X(z) = z X(lq) + X(lq + 1)
  int lq[2], l_norm[2];
  double l_coef[2];
  int nl;
  const void* xo[2];  // own-node terms (amplitude vectors)
  double o_coef[2];
  int no;
  void* out;
  int out_sq;
  const void* dotu;  // partial dot <dotu, out> per member (bilinear)
  void* partial;
  // k_tile only: out2 = the node-local species-diagonal terms at the nodes without flip partners (the nodes with partners
  // are left to k_rand_nodes): c2 X_q2 out (x2 null), or, for complex members, c2x X_q2x(z) x2~ + c2 X_q2(z) out with
  // the energy-linear pairs (q, q + 1) and x2~ = x2 / sq when x2norm
  void* out2;
  int q2;
  double c2;
  const void* x2;
  int q2x, x2norm;
  double c2x;
  // k_tile, complex members, the fused level of the edc chains (one pass fewer per level):
  // apart: partials of a_k = sum_i sq_i^2 [x~ (E + X(z)) x~ + x~ h1 - h1 (o h1)]_i into `partial`, at the nodes without
  //   flip partners (uo: the chain vector u, h1 = the pass's output, o h1 = its out2 or zero; a_e: the E term is present);
  // upd: the three-term update on the output, w = y - a u - b (s / s_prev) u_prev, the projection onto the symmetric
  //   subspace (proj) and the partials of <w, w> into `partial` (aeff, beff, scv, scpv: per member; act: activity)
  //   partial2: the partials of <u, y> before the update (the direct a_k, checked against the fused one by k_fin_b)
  int apart, a_e, upd, proj;
  void* partial2;
  const void* uo;
  const void* upv;
  const void *aeff, *beff, *scv, *scpv;
  const int* act;
  // the chain blocks (route ed3, EPI), O2 = 3 (the update pass of a block level): the partials of <tv, w> of the second
  // vector tv with the updated and projected w, into partial3 (c_k+1 = <tv, w> / b_k)
  const void* tvv;
  void* partial3;
  // fold (the fused level): 1 in the first pass, which stores g = h1 + (E + X(z)) x~ (route ed3: h1 + (z - et) x~ - u h1)
  // in place of h1 at the nodes without flip partners; 2 in the second pass, which leaves out its terms on uo and its
  // energy-linear terms there (their value comes with the xo input)
  int fold;
  // k_tile, real members, the pass that writes the output y (O2 = 0, no update): the stored output becomes
  //   y' = lcc_0 lcv_0 + ... + lcc_{lcn-1} lcv_{lcn-1} + y  (own rows, per-member coefficients; lcc_t null: 1), with the
  //   partial dot <dotu, y'>, in k_lincomb's operation order; and out3 = rib (y - ra u3 - rb up3) (a recomputed level in
  //   k_recompute's operation order; ra, rb, rib per member)
  int lcn;
  const void* lcv[3];
  const void* lcc[3];
  void* out3;
  const void *u3, *up3, *ra, *rb, *rib;
};

template <class T>
__device__ void block_reduce_member(T v, int mt, T* lds, T* partial_out) {
  // lanes with the same member (lane % mt) are summed over the node index within the wave, then over the waves
  const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x >> 6;
  for (int off = mt; off < WAVE; off <<= 1) v = v + shfl_xor(v, off);
  if (lane < mt) lds[wave * WAVE + lane] = v;
  __syncthreads();
  if (threadIdx.x < (unsigned)mt) {
    T s = lds[threadIdx.x];
    for (int w = 1; w < BLOCK / WAVE; ++w) s = s + lds[w * WAVE + threadIdx.x];
    *partial_out = s;
  }
}

// as block_reduce_member, adding the block's sum to the slot (the slot is owned by this block and lane)
template <class T>
__device__ void block_reduce_member_add(T v, int mt, T* lds, T* partial_out) {
  const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x >> 6;
  for (int off = mt; off < WAVE; off <<= 1) v = v + shfl_xor(v, off);
  if (lane < mt) lds[wave * WAVE + lane] = v;
  __syncthreads();
  if (threadIdx.x < (unsigned)mt) {
    T s = lds[threadIdx.x];
    for (int w = 1; w < BLOCK / WAVE; ++w) s = s + lds[w * WAVE + threadIdx.x];
    *partial_out = *partial_out + s;
  }
}

template <class T>
__global__ void __launch_bounds__(BLOCK) k_fused(OpDev op, MemDev md, FArgs fa) {
  extern __shared__ double lds_raw[];
  const int nblk = op.ntype * (op.nslot + 1) * 81;
  double* blk = lds_raw;
  int* pi = (int*)(blk + nblk);
  double* sg = (double*)(pi + ((op.nop * 9 + 1) & ~1));
  T* red = (T*)(sg + op.nop * 9);
  if (md.tact && !md.tact[blockIdx.y]) return;  // every member of this tile has left
  for (int q = threadIdx.x; q < nblk; q += BLOCK) blk[q] = op.blk[q];
  for (int q = threadIdx.x; q < op.nop * 9; q += BLOCK) { pi[q] = op.pi[q]; sg[q] = op.sg[q]; }
  __syncthreads();

  const int mt = md.mt, npw = WAVE / mt;
  const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x >> 6;
  const int m = blockIdx.y * mt + (lane & (mt - 1));
  const int i = (xcd_block(blockIdx.x, gridDim.x) * (BLOCK / WAVE) + wave) * npw + lane / mt;
  const bool live = (i < op.nn) && (m < md.nb);
  T dotv = zero<T>();
  if (live) {
    const uint32_t inf = op.info[i];
    const uint64_t chim = md.chimask[m];
    T acc[9];
    for (int l = 0; l < 9; ++l) acc[l] = zero<T>();
    if (fa.xs) {
      double dv[9];
      for (int l = 0; l < 9; ++l) dv[l] = md.dvec[(size_t)l * md.nbp + m];
      T sacc[9];
      for (int l = 0; l < 9; ++l) sacc[l] = zero<T>();
      s_accum<T>(op, blk, pi, sg, chim, i, (inf >> 12) & 255u, m, md.nbp, (const T*)fa.xs, fa.xs_norm, fa.s_din ? dv : nullptr, md.dperm,
                 sacc);
      for (int l = 0; l < 9; ++l) acc[l] = fma_s(fa.s_coef * (fa.s_dout ? dv[l] : 1.0), sacc[l], acc[l]);
    }
    for (int t = 0; t < fa.nr; ++t)
      rand_accum<T>(op, md, pi, sg, chim, i, inf, m, fa.qr[t], (const T*)fa.xr[t], fa.r_norm[t], fa.r_coef[t], acc);
    if (fa.nl > 0) lin_accum(op, md, pi, sg, chim, i, inf, m, fa.lq[0], (const T*)fa.xl[0], fa.l_norm[0], fa.l_coef[0], acc);
    if (fa.nl > 1) lin_accum(op, md, pi, sg, chim, i, inf, m, fa.lq[1], (const T*)fa.xl[1], fa.l_norm[1], fa.l_coef[1], acc);
    for (int t = 0; t < fa.no; ++t) {
      const T* xo = (const T*)fa.xo[t];
      for (int l = 0; l < 9; ++l) acc[l] = fma_s(fa.o_coef[t], ld(xo, i, l, md.nbp, m), acc[l]);
    }
    const double so = fa.out_sq ? op.sq[i] : 1.0;
    T* out = (T*)fa.out;
    const T* du = (const T*)fa.dotu;
    for (int l = 0; l < 9; ++l) {
      T y = so * acc[l];
      out[((size_t)i * 9 + l) * md.nbp + m] = y;
      if (du) dotv = dotv + bilin(ld(du, i, l, md.nbp, m), y);
    }
  }
  if (fa.dotu) {
    T* part = (T*)fa.partial;
    block_reduce_member<T>(dotv, mt, red, part + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * mt + threadIdx.x);
  }
}

// The fused operator pass with a small member tile MT (8 or 16; 64 / MT nodes per wave, one member per lane, real or
// complex members), every operation mapping each orbital onto itself (pi = identity) and NSLOT links per node.  The S pass
// reads each neighbour row (9 values of the lane's member) with raw buffer loads: the lane's offset (row of the
// neighbour, member) is one integer operation per link, the orbital offset a scalar register.  With one site type the
// bond block is uniform over the wave and its columns are scalar operands of the FMAs, in column order so that each
// gathered row's registers take the next link's row as soon as its FMAs are issued (the next slot's gathers are in flight
// during this slot's FMAs; the next column's block values and scale are requested one column ahead).  At 3 waves per
// SIMD (168 VGPRs) the headline level ran 10 % faster than with the row-order loop at 4 waves, and 3 % faster than this
// loop at 4 waves; 2 waves ran as fast as 3.  With several site types the same loop reads each block column from the node
// type's padded rows in LDS.  The scale of a gathered value, dd(l) chi(g) sg(g, l) / sq(r), comes
// from an LDS table and is shared by the real and imaginary parts.  The species-diagonal terms use the same loads and an
// LDS sign table; an energy-linear term X(z) = z X(a) + X(b) is one pass with a complex weight.  Member tiles of 8-16 keep one member slice
// of a 1-2 GB vector small enough to stay in the last-level caches while all nodes are swept.
// Row l of the lane's member at node offset voff: 16-byte (complex) or 8-byte (real) raw buffer load.
template <class T>
__device__ __forceinline__ T tile_row(__amdgpu_buffer_rsrc_t xr, uint32_t voff, uint32_t soff) {
  if constexpr (std::is_same<T, cplx>::value) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(xr, (int)voff, (int)soff, 0);
    return cplx{__builtin_bit_cast(double, (unsigned long long)v[0] | ((unsigned long long)v[1] << 32)),
                __builtin_bit_cast(double, (unsigned long long)v[2] | ((unsigned long long)v[3] << 32))};
  } else {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(xr, (int)voff, (int)soff, 0));
  }
}

// acc += coef X(x~) for the members of a tile (apply_rand): the node's own state and its flip partners, the member's local
// weights, the partner's sign chi(g) sg(g, l) from the LDS table cs (member x operation x orbital) and its row pi(g, l),
// raw buffer loads for the rows.  LIN = false: X = X_q with the real weights wt(q, class, l0, k, l).  LIN = true: the
// energy-dependent diagonal X(z) = z X(q) + X(q + 1) with the complex weight z wt(q) + wt(q + 1) (one pass, one accumulator).
// Frozen nodes: the frozen species' parameters on the node's own row.
template <class T, bool LIN>
__device__ __forceinline__ void rand_ct(const OpDev& op, const MemDev& md, const double* __restrict__ cs, int i, uint32_t inf,
                                        int m, uint32_t node16, uint32_t vs16, uint32_t coff, __amdgpu_buffer_rsrc_t xr,
                                        int xnorm, int q, double coef, T acc[9]) {
  const int nbp = md.nbp, nsm = op.nsmax;
  const int fs = frozen_slot(inf, md, m);
  const double isi = xnorm ? op.isq[i] : 1.0;
  const int qa = q, qb = LIN ? q + 1 : q;
  auto row = [&](uint32_t voff, int l) -> T { return tile_row<T>(xr, voff, l * vs16); };
  // acc += (coef (z wa + wb) or coef wa) xv
  auto add = [&](double wa, double wb, T xv, int l) {
    if constexpr (LIN && std::is_same<T, cplx>::value) {
      const cplx z = md.z[m];
      const cplx cw = cplx{coef * fma(z.re, wa, wb), coef * z.im * wa};
      acc[l] = acc[l] + cw * xv;
    } else {
      acc[l] = fma_s(coef * wa, xv, acc[l]);
    }
  };
  const uint32_t own_off = (uint32_t)i * node16 + coff;
  if (fs) {
    const double* fa_ = frozen_row(md, fs, qa) + m;
    const double* fb_ = frozen_row(md, fs, qb) + m;
#pragma unroll
    for (int l = 0; l < 9; ++l) {
      add(fa_[(size_t)l * nbp], LIN ? fb_[(size_t)l * nbp] : 0.0, isi * row(own_off, l), l);
    }
    return;
  }
  const int l0 = inf & 15u, kap = (inf >> 4) & 255u;
  const double* wa_ = md.wts + (((size_t)qa * op.ncls + kap) * nsm + l0) * nsm * 9 * nbp + m;
  const double* wb_ = md.wts + (((size_t)qb * op.ncls + kap) * nsm + l0) * nsm * 9 * nbp + m;
#pragma unroll
  for (int l = 0; l < 9; ++l) {
    add(wa_[((size_t)l0 * 9 + l) * nbp], LIN ? wb_[((size_t)l0 * 9 + l) * nbp] : 0.0, isi * row(own_off, l), l);
  }
  const int nsc = op.nsc[kap];
  const uint32_t* fl = op.flip + (size_t)i * nsm;
  for (int k = 0; k < nsc; ++k) {
    const uint32_t e = fl[k];
    if (k == l0 || e == 0u) continue;
    const int r = (int)(e >> 6) - 1, g = (int)(e & 63u);
    const double isr = xnorm ? op.isq[r] : 1.0;
    const double* __restrict__ cg = cs + g * 9;
    const int* __restrict__ pg = op.pi + g * 9;
    const uint32_t voff = (uint32_t)r * node16 + coff;
    // orbital l of the partner's image is its row pi(g, l) (the identity for an orbital-preserving group)
#pragma unroll
    for (int l = 0; l < 9; ++l)
      add(wa_[((size_t)k * 9 + l) * nbp], LIN ? wb_[((size_t)k * 9 + l) * nbp] : 0.0,
          (isr * cg[l]) * tile_row<T>(xr, voff + (uint32_t)pg[l] * vs16, 0u), l);
  }
}

// streaming (non-temporal) stores and loads of values a pass writes or reads once: those lines need not take L2 room from
// the gathered rows (measured on the headline level: 143.5 -> 134.9 ms per 30 levels, the coefficients unchanged)
typedef double nt_d2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void nt_store(cplx* p, cplx v) { __builtin_nontemporal_store(nt_d2{v.re, v.im}, (nt_d2*)p); }
__device__ __forceinline__ void nt_store(double* p, double v) { __builtin_nontemporal_store(v, p); }
// an output value: streamed when a wave's row segment fills whole 128-byte lines, else a plain store
template <bool NT, class T>
__device__ __forceinline__ void out_store(T* p, T v) {
  if constexpr (NT) nt_store(p, v);
  else *p = v;
}
__device__ __forceinline__ cplx nt_load(const cplx* p) {
  const nt_d2 v = __builtin_nontemporal_load((const nt_d2*)p);
  return cplx{v.x, v.y};
}
__device__ __forceinline__ double nt_load(const double* p) { return __builtin_nontemporal_load(p); }

// O2: the out2 form compiled in (0 none, 1 c2 X_q2 out, 2 the two energy-linear terms of route ed3's t pass)
// PI: every operation of the group maps each orbital onto itself (pi = identity, e.g. D2h); otherwise (e.g. the D4h of
// class_ops) orbital l2 of a gathered image is row pi(g, l2) of the neighbour, whose nine 4-bit row indices per operation
// come packed from LDS, and the gather scale carries d of that row.
template <class T, int MT, int NSLOT, int O2, bool PI>
__global__ void __launch_bounds__(BLOCK, (O2 == 3 ? 2 : 3)) k_tile(OpDev op, MemDev md, FArgs fa, const double* __restrict__ sblk) {
  extern __shared__ double bls[];
  __shared__ T red[BLOCK];
  __shared__ unsigned long long pik[PI ? 1 : 64];
  if constexpr (!PI) {
    for (int g = threadIdx.x; g < op.nop; g += BLOCK) {
      unsigned long long pk = 0ull;
      for (int l = 0; l < 9; ++l) pk |= (unsigned long long)op.pi[g * 9 + l] << (4 * l);
      pik[g] = pk;
    }
  }
  constexpr uint32_t ES = sizeof(T);
  if (md.tact && !md.tact[blockIdx.y]) return;  // every member of this tile has left
  // LDS: the padded bond blocks, then the gather scales e[member][g][l] = dd(l) chi(g) sg(g, l) of the tile's members
  const int nblk = op.ntype * (op.nslot + 1) * 90;
  double* etab = bls + nblk;
  double* cst = etab + MT * (op.nop + 1) * 9;  // chi(g) sg(g, l) per member of the tile
  for (int q = threadIdx.x; q < MT * op.nop * 9; q += BLOCK) {
    const int ml = q / (op.nop * 9), g = (q / 9) % op.nop, l = q % 9;
    const int mm = min(blockIdx.y * MT + ml, md.nb - 1);
    cst[q] = (((md.chimask[mm] >> g) & 1ull) ? -1.0 : 1.0) * op.sg[g * 9 + l];
  }
  if (!fa.xs) __syncthreads();
  if (fa.xs) {
    for (int q = threadIdx.x; q < nblk; q += BLOCK) bls[q] = op.blkp[q];
    // row g = nop: the on-site term (no operation)
    const int ng = op.nop + 1, ne = MT * ng * 9;
    for (int q = threadIdx.x; q < ne; q += BLOCK) {
      const int ml = q / (ng * 9), g = (q / 9) % ng, l = q % 9;
      const int mm = min(blockIdx.y * MT + ml, md.nb - 1);
      const double d = !fa.s_din ? 1.0
                       : (PI || g == op.nop) ? md.dvec[(size_t)l * md.nbp + mm]
                                             : md.dperm[((size_t)g * 9 + l) * md.nbp + mm];
      const double chi = (g < op.nop && ((md.chimask[mm] >> g) & 1ull)) ? -1.0 : 1.0;
      etab[q] = d * chi * (g < op.nop ? op.sg[g * 9 + l] : 1.0);
    }
    __syncthreads();
  }
  constexpr int npw = WAVE / MT;
  const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x >> 6;
  const int ml = lane & (MT - 1);
  const int m = blockIdx.y * MT + ml;
  const int i = (xcd_block(blockIdx.x, gridDim.x) * (BLOCK / WAVE) + wave) * npw + lane / MT;
  const bool live = (i < op.nn) && (m < md.nb);
  const int nbp = md.nbp;
  T dotv = zero<T>(), dotv2 = zero<T>(), dotv3 = zero<T>();
  if (live) {
    const uint32_t inf = op.info[i];
    T acc[9];
    for (int l = 0; l < 9; ++l) acc[l] = zero<T>();
    if (fa.xs) {
      const uint32_t vs16 = (uint32_t)nbp * ES, node16 = 9u * vs16, coff = (uint32_t)m * ES;
      const __amdgpu_buffer_rsrc_t xr = __builtin_amdgcn_make_buffer_rsrc((void*)fa.xs, (short)0,
                                                                          (int)((uint32_t)(op.nn + 1) * node16), 0x00020000);
      const double* __restrict__ bl = bls + (size_t)((inf >> 12) & 255u) * (NSLOT + 1) * 90;
      const double* __restrict__ em = etab + (size_t)ml * (op.nop + 1) * 9;
      const uint32_t* __restrict__ hp = op.hop + (size_t)i * NSLOT;
      // Slot k = 0: the node itself with the on-site block; k > 0: link k - 1.  Column order: gathered row l2 of slot k is
      // dead after its 9 FMAs, and the same registers take row l2 of slot k + 1 at once, so the next slot's gathers are
      // in flight while this slot's FMAs run, in the registers of one slot.  Each acc[l] takes its terms in the order
      // (k, l2) of a row-order sum.  The block value B_k[l][l2]: with one site type uniform over the wave, from the
      // transposed copy through scalar loads (sblk holds column l2 of slot k at k * 90 + l2 * 10); otherwise from the
      // node type's padded rows in LDS.
      auto run = [&](auto onetype) {
        auto bv = [&](int k, int l2, int l) -> double {
          if constexpr (decltype(onetype)::value) return sblk[(size_t)k * 90 + l2 * 10 + l];
          else return bl[k * 90 + l * 10 + l2];
        };
        // per slot, in the order of the loop body: 1/sq of the next slot's node (raw), the link after it, then the rows
        int g = op.nop;
        double q = op.isq[i];
        uint32_t en = hp[0];  // the link of slot k + 1, read one slot ahead
        uint32_t voff = (uint32_t)i * node16 + coff;
        unsigned long long pm = 0ull;  // !PI: the packed rows pi(g, .) of the slot whose rows are being gathered
        T v[9];
#pragma unroll
        for (int l = 0; l < 9; ++l) {
          v[l] = tile_row<T>(xr, voff, l * vs16);
          __builtin_amdgcn_sched_barrier(0);
        }
        // the block column and the gather scale of the next column are requested before this column's FMAs
        double bc[9], ec;
#pragma unroll
        for (int l = 0; l < 9; ++l) bc[l] = bv(0, 0, l);
        ec = em[op.nop * 9];
        auto slot = [&](int k, auto more) {
          const double* __restrict__ eg = em + g * 9;
          const double sck = fa.xs_norm ? q : 1.0;
          if constexpr (decltype(more)::value) {
            const int r = (int)(en >> 6) - 1;  // links of no neighbour hold the zero padding node nn
            g = (int)(en & 63u);
            voff = (uint32_t)r * node16 + coff;
            q = op.isq[r];
            en = hp[k + 1 < NSLOT ? k + 1 : k];
            if constexpr (!PI) pm = pik[g];
          }
#pragma unroll
          for (int l2 = 0; l2 < 9; ++l2) {
            double bn[9], egn = 0.0;
            if (l2 < 8) {
#pragma unroll
              for (int l = 0; l < 9; ++l) bn[l] = bv(k, l2 + 1, l);
              egn = eg[l2 + 1];
            } else if constexpr (decltype(more)::value) {
#pragma unroll
              for (int l = 0; l < 9; ++l) bn[l] = bv(k + 1, 0, l);
              egn = em[g * 9];
            }
            const T x = (sck * ec) * v[l2];
#pragma unroll
            for (int l = 0; l < 9; ++l) acc[l] = fma_s(bc[l], x, acc[l]);
            if constexpr (decltype(more)::value) {
              if constexpr (PI) v[l2] = tile_row<T>(xr, voff, l2 * vs16);
              else v[l2] = tile_row<T>(xr, voff + (uint32_t)((pm >> (4 * l2)) & 15ull) * vs16, 0u);
            }
            if (l2 < 8 || decltype(more)::value) {
#pragma unroll
              for (int l = 0; l < 9; ++l) bc[l] = bn[l];
              ec = egn;
            }
            __builtin_amdgcn_sched_barrier(0);
          }
        };
#pragma unroll 1
        for (int k = 0; k < NSLOT; ++k) slot(k, std::true_type{});
        slot(NSLOT, std::false_type{});
      };
      if (op.ntype == 1) run(std::true_type{});
      else run(std::false_type{});
      for (int l = 0; l < 9; ++l) acc[l] = (fa.s_coef * (fa.s_dout ? md.dvec[(size_t)l * nbp + m] : 1.0)) * acc[l];
    }
    {
      const uint32_t vs16 = (uint32_t)nbp * ES, node16 = 9u * vs16, coff = (uint32_t)m * ES;
      const int nrec = (int)((uint32_t)(op.nn + 1) * node16);
      const double* __restrict__ cs = cst + (size_t)ml * op.nop * 9;
      // fold = 2 (second pass of the fused level): the species-diagonal terms on the chain vector (fa.uo) and the
      // energy-linear ones at nodes without flip partners are in the xo input already (g of the first pass)
      const bool skip_u = fa.fold == 2 && !(inf & INFO_PARTNERS);
      for (int t = 0; t < fa.nr; ++t) {
        if (skip_u && fa.xr[t] == fa.uo) continue;
        const __amdgpu_buffer_rsrc_t rr = __builtin_amdgcn_make_buffer_rsrc((void*)fa.xr[t], (short)0, nrec, 0x00020000);
        rand_ct<T, false>(op, md, cs, i, inf, m, node16, vs16, coff, rr, fa.r_norm[t], fa.qr[t], fa.r_coef[t], acc);
      }
      if constexpr (std::is_same<T, cplx>::value) {
        for (int t = 0; t < fa.nl; ++t) {
          if (skip_u) continue;  // the energy-linear terms (on u, and route ed3's on h1) are in g
          // X(z) x~ = z X(a) x~ + X(b) x~ in one pass over x
          const __amdgpu_buffer_rsrc_t rl = __builtin_amdgcn_make_buffer_rsrc((void*)fa.xl[t], (short)0, nrec, 0x00020000);
          rand_ct<T, true>(op, md, cs, i, inf, m, node16, vs16, coff, rl, fa.l_norm[t], fa.lq[t], fa.l_coef[t], acc);
        }
      }
    }
    for (int t = 0; t < fa.no; ++t) {
      const T* xo = (const T*)fa.xo[t];
      for (int l = 0; l < 9; ++l) acc[l] = fma_s(fa.o_coef[t], nt_load(&xo[((size_t)i * 9 + l) * nbp + m]), acc[l]);
    }
    const double so = fa.out_sq ? op.sq[i] : 1.0;
    T* out = (T*)fa.out;
    const T* du = (const T*)fa.dotu;
    constexpr bool NTO = MT * sizeof(T) >= 128;
    // the own-state weights of the node for parameter slot q (the frozen species' row on a frozen site)
    auto wrow = [&](int q) -> const double* {
      const int fs = frozen_slot(inf, md, m);
      const int l0 = inf & 15u, kap = (inf >> 4) & 255u, nsm = op.nsmax;
      return fs ? frozen_row(md, fs, q) + m : md.wts + ((((size_t)q * op.ncls + kap) * nsm + l0) * nsm + l0) * 9 * nbp + m;
    };
    // fold (the fused edc level, O2 = 1): at a node without flip partners the stored output is g = h1 + (E + X(z)) x~,
    // the node-local terms the second pass would otherwise read x~ and the weights for, and t1 = c2 X_q2 h1 is kept in t1r
    // from the a_k terms below for the out2 store
    T t1r[O2 == 1 ? 9 : 1];
    const bool folded = O2 == 1 && fa.fold == 1 && fa.apart && !(inf & INFO_PARTNERS);
    if constexpr (std::is_same<T, cplx>::value && O2 < 2) {
      // a_k at a node without flip partners: sq^2 [x~ (E + X(z)) x~ + x~ h1 - h1 t1] with h1 = y (out_sq = 0) and
      // t1 = c2 X_q2 h1 (O2 = 1) or zero; (E + X(z)) and X_q2 are node-local there (route ed3: in the out2 loop below)
      if (fa.apart && !(inf & INFO_PARTNERS)) {
        const cplx z = md.z[m];
        const double* we = fa.a_e ? wrow(QE) : nullptr;
        const double *wpa = wrow(QPA), *wpb = wrow(QPB), *wo = O2 == 1 ? wrow(fa.q2) : nullptr;
        const double isi = op.isq[i], sq2 = op.sq[i] * op.sq[i];
        const cplx* uo = (const cplx*)fa.uo;
        cplx pa = zero<cplx>();
        for (int l = 0; l < 9; ++l) {
          const cplx xt = isi * uo[((size_t)i * 9 + l) * nbp + m];
          const double wal = wpa[(size_t)l * nbp];
          const cplx cw = cplx{fma(z.re, wal, wpb[(size_t)l * nbp]) + (we ? we[(size_t)l * nbp] : 0.0), z.im * wal};
          const cplx h = so * acc[l];
          const cplx xw = cw * xt;
          cplx term = bilin(xt, xw) + bilin(xt, h);
          if constexpr (O2 == 1) {
            const cplx t1 = fma_s(fa.c2 * wo[(size_t)l * nbp], 1.0 * h, zero<T>());
            term = term - bilin(h, t1);
            if (folded) {
              t1r[l] = t1;
              acc[l] = h + xw;  // so = 1 in this pass
            }
          }
          pa = pa + term;
        }
        dotv = sq2 * pa;
      }
    }
    if constexpr (O2 == 0 || O2 == 3) {
      bool done = false;
      if constexpr (std::is_same<T, cplx>::value) {
        if (fa.upd) {
          // w = y - a u - b (s_k / s_{k-1}) u_prev (lagged normalisation; zero for a member that has left), projected
          // onto the symmetric subspace (every operation keeps each orbital: a factor per orbital), then <w, w>
          const cplx a = ((const cplx*)fa.aeff)[m];
          const cplx b = bilin(((const cplx*)fa.beff)[m], bilin(((const cplx*)fa.scv)[m], inv(((const cplx*)fa.scpv)[m])));
          const bool off = !fa.act[m];
          double f[9];
          int ns = 1;
          if (fa.proj) {
            const int s0 = op.stab_ptr[i], s1 = op.stab_ptr[i + 1];
            ns = s1 - s0;
            if (PI && ns > 1) {
              const double* __restrict__ cs = cst + (size_t)ml * op.nop * 9;
              for (int l = 0; l < 9; ++l) f[l] = 0.0;
              for (int j = s0; j < s1; ++j) {
                const int g = op.stab[j];
                for (int l = 0; l < 9; ++l) f[l] += cs[g * 9 + l];
              }
            }
          }
          const cplx* uo = (const cplx*)fa.uo;
          const cplx* up = (const cplx*)fa.upv;
          cplx ww = zero<cplx>(), uy = zero<cplx>(), tw = zero<cplx>();
          // O2 = 3: the second vector's rows (with orbital-preserving operations requested with the other operands)
          const cplx* tvr = (const cplx*)fa.tvv;
          cplx tq[O2 == 3 && PI ? 9 : 1];
          if constexpr (O2 == 3 && PI) {
#pragma unroll
            for (int l = 0; l < 9; ++l) tq[l] = tvr[((size_t)i * 9 + l) * nbp + m];
          }
          if constexpr (PI) {
            // the nine u and u_prev values are requested before the first store: the compiler cannot tell the output
            // from them, and a load behind a store would wait for the store as well (one vector-memory counter).
            // Measured on the headline level: the second operator pass 1.97 -> 1.88 ms (writing the nine values to the
            // accumulators first and storing after, 1.93 ms)
            cplx uq[9], upq[9];
#pragma unroll
            for (int l = 0; l < 9; ++l) {
              const size_t q = ((size_t)i * 9 + l) * nbp + m;
              uq[l] = uo[q];
              upq[l] = NTO ? nt_load(&up[q]) : up[q];
            }
#pragma unroll
            for (int l = 0; l < 9; ++l) {
              const size_t q = ((size_t)i * 9 + l) * nbp + m;
              const cplx y = so * acc[l];
              uy = uy + bilin(uq[l], y);
              cplx w = off ? zero<cplx>() : y - bilin(a, uq[l]) - bilin(b, upq[l]);
              if (ns > 1) w = (f[l] / (double)ns) * w;
              out_store<NTO>(&out[q], w);
              ww = ww + bilin(w, w);
              if constexpr (O2 == 3) tw = tw + bilin(tq[l], w);
            }
          } else {
            // (P w)_l = sum_{g in S} chi(g) sg(g, l) w_pi(g, l) / |S|: the nine w of the node first (in acc), then each
            // w_pi from the registers (a run-time index into a register array would live in scratch memory)
#pragma unroll
            for (int l = 0; l < 9; ++l) {
              const size_t q = ((size_t)i * 9 + l) * nbp + m;
              const cplx y = so * acc[l], uq = uo[q];
              uy = uy + bilin(uq, y);
              acc[l] = off ? zero<cplx>() : y - bilin(a, uq) - bilin(b, NTO ? nt_load(&up[q]) : up[q]);
            }
            if (ns > 1) {
              // w_p as a sum of the nine values weighted 1 or 0 (a select chain or an index becomes an indexed load from
              // scratch memory); exact for finite values
              const double* __restrict__ cs = cst + (size_t)ml * op.nop * 9;
              const int s0 = op.stab_ptr[i];
#pragma unroll 1
              for (int l = 0; l < 9; ++l) {
                cplx sacc = zero<cplx>();
                for (int j = s0; j < s0 + ns; ++j) {
                  const int g = op.stab[j], p = op.pi[g * 9 + l];
                  cplx wv = zero<cplx>();
#pragma unroll
                  for (int l2 = 0; l2 < 9; ++l2) wv = fma_s(p == l2 ? 1.0 : 0.0, acc[l2], wv);
                  sacc = sacc + cs[g * 9 + l] * wv;
                }
                const cplx w = (1.0 / (double)ns) * sacc;
                out_store<NTO>(&out[((size_t)i * 9 + l) * nbp + m], w);
                ww = ww + bilin(w, w);
                if constexpr (O2 == 3) tw = tw + bilin(tvr[((size_t)i * 9 + l) * nbp + m], w);
              }
            } else {
#pragma unroll
              for (int l = 0; l < 9; ++l) {
                out_store<NTO>(&out[((size_t)i * 9 + l) * nbp + m], acc[l]);
                ww = ww + bilin(acc[l], acc[l]);
                if constexpr (O2 == 3) tw = tw + bilin(tvr[((size_t)i * 9 + l) * nbp + m], acc[l]);
              }
            }
          }
          dotv = ww;
          dotv2 = uy;
          dotv3 = tw;
          done = true;
        }
      }
      if constexpr (std::is_same<T, double>::value) {
        if (fa.lcn > 0 || fa.out3) {
          // the adjoint's combinations on the output in registers (the separate passes k_lincomb and k_recompute)
          double lc[3];
          for (int t = 0; t < 3; ++t) lc[t] = t < fa.lcn ? (fa.lcc[t] ? ((const double*)fa.lcc[t])[m] : 1.0) : 0.0;
          const double ra = fa.out3 ? ((const double*)fa.ra)[m] : 0.0, rb = fa.out3 ? ((const double*)fa.rb)[m] : 0.0,
                       rib = fa.out3 ? ((const double*)fa.rib)[m] : 0.0;
          // every row this epilogue reads is requested before the first store (the stores may alias them for the
          // compiler, which would otherwise wait for each load in turn)
          const size_t q0 = (size_t)i * 9 * nbp + m;
          double z[9], p[9];
          for (int l = 0; l < 9; ++l) {
            z[l] = 0.0;
            p[l] = 0.0;
          }
          if (fa.lcn > 0) {
            for (int t = 0; t < fa.lcn; ++t) {
              const double* __restrict__ v = (const double*)fa.lcv[t] + q0;
#pragma unroll
              for (int l = 0; l < 9; ++l) z[l] = fma(lc[t], nt_load(&v[(size_t)l * nbp]), z[l]);
            }
          } else if (fa.out3) {
            const double* __restrict__ u3 = (const double*)fa.u3 + q0;
            const double* __restrict__ up3 = (const double*)fa.up3 + q0;
#pragma unroll
            for (int l = 0; l < 9; ++l) {
              z[l] = u3[(size_t)l * nbp];
              p[l] = nt_load(&up3[(size_t)l * nbp]);
            }
          }
          double dv[9];
#pragma unroll
          for (int l = 0; l < 9; ++l) dv[l] = du ? du[q0 + (size_t)l * nbp] : 0.0;
#pragma unroll
          for (int l = 0; l < 9; ++l) {
            const size_t q = q0 + (size_t)l * nbp;
            const double y = so * acc[l];
            const double o = fa.lcn > 0 ? fma(1.0, y, z[l]) : y;
            out_store<NTO>(&out[q], o);
            if (du) dotv = dotv + bilin(dv[l], o);
            if (fa.out3) {
              const double xv = y - bilin(ra, z[l]) - bilin(rb, p[l]);
              out_store<NTO>(&((double*)fa.out3)[q], bilin(rib, xv));
            }
          }
          done = true;
        }
      }
      if (!done) {
        // the dot's rows before the first store (see the update above)
        T dv[9];
#pragma unroll
        for (int l = 0; l < 9; ++l) dv[l] = du ? ld(du, i, l, nbp, m) : zero<T>();
#pragma unroll
        for (int l = 0; l < 9; ++l) {
          const T y = so * acc[l];
          out_store<NTO>(&out[((size_t)i * 9 + l) * nbp + m], y);
          if (du) dotv = dotv + bilin(dv[l], y);
        }
      }
    } else {
      // out2: the own-state terms of rand_ct on the output row (and on x2), in rand_ct's operation order, no partner rows
      T* out2 = !(inf & INFO_PARTNERS) ? (T*)fa.out2 : nullptr;
      const double *wa = nullptr, *wb = nullptr, *xa = nullptr, *xb = nullptr, *pa_ = nullptr, *pb_ = nullptr;
      if (out2) {
        wa = wrow(fa.q2);
        if constexpr (O2 == 2) {
          wb = wrow(fa.q2 + 1);
          xa = wrow(fa.q2x);
          xb = wrow(fa.q2x + 1);
          if (fa.apart) {
            pa_ = wrow(QPA);
            pb_ = wrow(QPB);
          }
        }
      }
      // route ed3's a_k at a node without flip partners (O2 = 2, apart; z - K' of ed3_kmul, x~ u h1 = h1 u x~ and
      // x~ h'(t) = h1 t on the symmetric subspace): sq^2 [x~ (z - et) x~ + x~ h1 - h1 (u x~) - h1 t], t = out2 (c2x = 1,
      // x~ = the normalised x2, the chain vector)
      // fold (O2 = 2, the fused level of route ed3): at a node without flip partners the stored output is
      // g = h1 + (z - et) x~ - u h1, the node-local energy-linear terms of the second pass
      cplx pa3 = zero<cplx>();
      for (int l = 0; l < 9; ++l) {
        const T y = so * acc[l];
        if constexpr (O2 != 2) out_store<(MT * sizeof(T) >= 128)>(&out[((size_t)i * 9 + l) * nbp + m], y);
        if (du) dotv = dotv + bilin(ld(du, i, l, nbp, m), y);
        if (out2) {
          if constexpr (O2 == 1) {
            out_store<(MT * sizeof(T) >= 128)>(&out2[((size_t)i * 9 + l) * nbp + m],
                                               folded ? t1r[l] : fma_s(fa.c2 * wa[(size_t)l * nbp], 1.0 * y, zero<T>()));
          } else if constexpr (std::is_same<T, cplx>::value) {
            const cplx z = md.z[m];
            const double sx = fa.x2norm ? op.isq[i] : 1.0;
            const cplx xv = sx * ((const cplx*)fa.x2)[((size_t)i * 9 + l) * nbp + m];
            const double xal = xa[(size_t)l * nbp], wal = wa[(size_t)l * nbp];
            const cplx cx = cplx{fa.c2x * fma(z.re, xal, xb[(size_t)l * nbp]), fa.c2x * z.im * xal};
            const cplx cy = cplx{fa.c2 * fma(z.re, wal, wb[(size_t)l * nbp]), fa.c2 * z.im * wal};
            const cplx tv = (zero<cplx>() + cx * xv) + cy * (1.0 * y);
            out_store<(MT * sizeof(T) >= 128)>(&out2[((size_t)i * 9 + l) * nbp + m], tv);
            cplx yo = y;
            if (pa_) {
              const double pal = pa_[(size_t)l * nbp];
              const cplx cw = cplx{fma(z.re, pal, pb_[(size_t)l * nbp]), z.im * pal};
              const cplx uz = cplx{fma(z.re, xal, xb[(size_t)l * nbp]), z.im * xal};
              pa3 = pa3 + (((bilin(xv, cw * xv) + bilin(xv, y)) - bilin(y, uz * xv)) - bilin(y, tv));
              if (fa.fold == 1) yo = (y + cw * xv) - uz * y;
            }
            out_store<(MT * sizeof(T) >= 128)>(&out[((size_t)i * 9 + l) * nbp + m], yo);
          } else {
            out_store<(MT * sizeof(T) >= 128)>(&out[((size_t)i * 9 + l) * nbp + m], y);
          }
        } else if constexpr (O2 == 2) {
          out_store<(MT * sizeof(T) >= 128)>(&out[((size_t)i * 9 + l) * nbp + m], y);
        }
      }
      if constexpr (O2 == 2 && std::is_same<T, cplx>::value) {
        if (pa_) dotv = (op.sq[i] * op.sq[i]) * pa3;
      }
    }
  }
  if (fa.dotu || fa.apart || fa.upd) {
    T* part = (T*)fa.partial;
    block_reduce_member<T>(dotv, MT, red, part + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * MT + threadIdx.x);
  }
  if (fa.partial2) {
    __syncthreads();
    T* part = (T*)fa.partial2;
    block_reduce_member<T>(dotv2, MT, red, part + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * MT + threadIdx.x);
  }
  if constexpr (O2 == 3) {
    __syncthreads();
    T* part = (T*)fa.partial3;
    block_reduce_member<T>(dotv3, MT, red, part + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * MT + threadIdx.x);
  }
}

// The species-diagonal term y = coef X_q x (rand_ct without normalisation) at the listed nodes only: the nodes with flip
// partners, whose rows k_tile's out2 leaves out.  One thread per (listed node, member), members fastest.  Same operation
// order as rand_ct, so the rows equal those of a full pass.
template <class T, class F>
__device__ inline void rand_node_rows(const OpDev& op, const MemDev& md, int i, uint32_t inf, int m, F xrow, int q,
                                      double coef, int xnorm, bool lin, T acc[9]) {
  // rand_ct for one (node, member), the rows of x from xrow(node, row) (acc += coef X_q x; lin: coef (z X_q + X_q+1) x,
  // complex members)
  const int nbp = md.nbp, nsm = op.nsmax;
  const int fs = frozen_slot(inf, md, m);
  const double isi = xnorm ? op.isq[i] : 1.0;
  auto add = [&](double wa, double wb, T xv, int l) {
    if constexpr (std::is_same<T, cplx>::value) {
      if (lin) {
        const cplx z = md.z[m];
        const cplx cw = cplx{coef * fma(z.re, wa, wb), coef * z.im * wa};
        acc[l] = acc[l] + cw * xv;
        return;
      }
    }
    acc[l] = fma_s(coef * wa, xv, acc[l]);
  };
  const int qb = lin ? q + 1 : q;
  if (fs) {
    const double* fa_ = frozen_row(md, fs, q) + m;
    const double* fb_ = frozen_row(md, fs, qb) + m;
    for (int l = 0; l < 9; ++l) add(fa_[(size_t)l * nbp], lin ? fb_[(size_t)l * nbp] : 0.0, isi * xrow(i, l), l);
    return;
  }
  const int l0 = inf & 15u, kap = (inf >> 4) & 255u;
  const double* wa_ = md.wts + (((size_t)q * op.ncls + kap) * nsm + l0) * nsm * 9 * nbp + m;
  const double* wb_ = md.wts + (((size_t)qb * op.ncls + kap) * nsm + l0) * nsm * 9 * nbp + m;
  for (int l = 0; l < 9; ++l)
    add(wa_[((size_t)l0 * 9 + l) * nbp], lin ? wb_[((size_t)l0 * 9 + l) * nbp] : 0.0, isi * xrow(i, l), l);
  const uint64_t chim = md.chimask[m];
  const int nsc = op.nsc[kap];
  const uint32_t* fl = op.flip + (size_t)i * nsm;
  for (int k = 0; k < nsc; ++k) {
    const uint32_t e = fl[k];
    if (k == l0 || e == 0u) continue;
    const int r = (int)(e >> 6) - 1, g = (int)(e & 63u);
    const double isr = xnorm ? op.isq[r] : 1.0;
    const double chi = ((chim >> g) & 1ull) ? -1.0 : 1.0;
    for (int l = 0; l < 9; ++l)
      add(wa_[((size_t)k * 9 + l) * nbp], lin ? wb_[((size_t)k * 9 + l) * nbp] : 0.0,
          (isr * (chi * op.sg[g * 9 + l])) * xrow(r, op.pi[g * 9 + l]), l);
  }
}

// rand_ct for one (node, member) with plain loads of x
template <class T>
__device__ inline void rand_node_term(const OpDev& op, const MemDev& md, int i, uint32_t inf, int m, const T* __restrict__ x,
                                      int q, double coef, int xnorm, bool lin, T acc[9]) {
  const int nbp = md.nbp;
  rand_node_rows<T>(op, md, i, inf, m, [&](int r, int l) -> T { return x[((size_t)r * 9 + l) * nbp + m]; }, q, coef, xnorm,
                    lin, acc);
}

// y = terms at the listed nodes: c2x X_q2x x2 (when x2) + c2 X_q2 x, the form of k_tile's out2 (O2 = 1: x2 null,
// plain weights; O2 = 2: energy-linear pairs)
template <class T>
__global__ void k_rand_nodes(OpDev op, MemDev md, const int* __restrict__ list, int nlist, const T* __restrict__ x,
                             T* __restrict__ y, int q, double coef, const T* __restrict__ x2, int q2x, int x2norm,
                             double c2x) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int nbp = md.nbp;
  const int j = (int)(t / nbp), m = (int)(t % nbp);
  if (j >= nlist || m >= md.nb) return;
  if (md.tact && !md.tact[m / md.mt]) return;
  const int i = list[j];
  const uint32_t inf = op.info[i];
  T acc[9];
  for (int l = 0; l < 9; ++l) acc[l] = zero<T>();
  if (x2) rand_node_term<T>(op, md, i, inf, m, x2, q2x, c2x, x2norm, true, acc);
  rand_node_term<T>(op, md, i, inf, m, x, q, coef, 0, x2 != nullptr, acc);
  for (int l = 0; l < 9; ++l) y[((size_t)i * 9 + l) * nbp + m] = acc[l];
}

// The listed nodes (those with flip partners) of the fused edc level, one thread per (listed node, member): t1 = X_o h1
// (when t1 is non-null, as k_rand_nodes) and the node's term of a_k, sq^2 [x~ (E + X(z)) x~ + x~ h1 - h1 t1] with the
// partner rows (lpart: nlist x nbp, summed by k_fin_a after the block partials of the tile pass).  Route ed3 (ed3 != 0):
// t1 = t = u x~ - Qt h1 (k_rand_nodes' order: the u term first) and sq^2 [x~ (z - et) x~ + x~ h1 - h1 (u x~) - h1 t].
__global__ void __launch_bounds__(256) k_list_terms(OpDev op, MemDev md, const int* __restrict__ list, int nlist,
                                                     const cplx* __restrict__ h1, cplx* __restrict__ t1,
                                                     const cplx* __restrict__ x, int a_e, cplx* __restrict__ lpart,
                                                     int ed3) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int nbp = md.nbp;
  const int j = (int)(t / nbp), m = (int)(t % nbp);
  if (j >= nlist || m >= md.nb) return;
  if (md.tact && !md.tact[m / md.mt]) return;
  const int i = list[j];
  const uint32_t inf = op.info[i];
  cplx to[9], ex[9], ux[9];
  for (int l = 0; l < 9; ++l) { to[l] = zero<cplx>(); ex[l] = zero<cplx>(); ux[l] = zero<cplx>(); }
  // every load before the stores at the end (a load behind a store waits for it as well)
  if (ed3) {
    rand_node_term<cplx>(op, md, i, inf, m, x, QUA, 1.0, 1, true, to);
    for (int l = 0; l < 9; ++l) ux[l] = to[l];
    rand_node_term<cplx>(op, md, i, inf, m, h1, QQA, -1.0, 0, true, to);
  } else if (t1) {
    rand_node_term<cplx>(op, md, i, inf, m, h1, QO, 1.0, 0, false, to);
  }
  if (a_e) rand_node_term<cplx>(op, md, i, inf, m, x, QE, 1.0, 1, false, ex);
  rand_node_term<cplx>(op, md, i, inf, m, x, QPA, 1.0, 1, true, ex);
  const double isi = op.isq[i], sq2 = op.sq[i] * op.sq[i];
  cplx pa = zero<cplx>();
  for (int l = 0; l < 9; ++l) {
    const size_t q = ((size_t)i * 9 + l) * nbp + m;
    const cplx xt = isi * x[q], h = h1[q];
    cplx term = (bilin(xt, ex[l]) + bilin(xt, h)) - bilin(h, to[l]);
    if (ed3) term = term - bilin(h, ux[l]);
    pa = pa + term;
  }
  if (t1)
    for (int l = 0; l < 9; ++l) t1[((size_t)i * 9 + l) * nbp + m] = to[l];
  lpart[(size_t)j * nbp + m] = sq2 * pa;
}

// The same fused operator pass with one node per wave and 64 members per wave (member tile 64): the node's links, weights
// table offsets, the orbital tables of each link's operation and the bond blocks are wave-uniform and come through the
// scalar cache; only the member amplitudes are per lane, each gather a uniform row address plus the lane's offset.
// NSLOT > 0: the slot loop is unrolled at compile time (a missing link points at the all-zero padding node nn, so the
// loop has no branch and the loads of later slots can be issued early).  PIID: every operation of the group maps each
// orbital onto itself (pi = identity, e.g. D2h), so the member's d enters at the orbital index itself.
// The S pass of one node for one real column per lane when every operation maps each orbital onto itself (pi = identity):
//   acc += sum_k f_k B_k (dd . sg_k . x_k),   f_k = chi(g_k) / sq(r_k) (xnorm) or chi(g_k),
// over the on-site block (k = 0) and the NSLOT links.  The neighbour rows are raw buffer loads: a uniform byte offset
// (row of the neighbour, orbital) on the scalar unit plus the lane's column offset, so no vector instruction is spent on
// addresses; the loads of link k + 1 are issued before the block product of link k.  The bond blocks of the node's type
// come from LDS (rows padded to 10 doubles), one row at a time next to its nine FMAs (a scheduling barrier per row keeps
// the block out of the registers).  The orbital signs sg_k are sign-bit flips with a uniform mask; chi(g_k) / sq(r_k) enters
// once per orbital.  Per link and column: 9 loads, 81 FMA, 18 multiplications, 9 sign flips.
// launch_fused takes this path only for vectors below 4 GiB (buffer_ok: 32-bit buffer offsets and record counts).
template <int NSLOT>
__device__ __forceinline__ void s_link_load(__amdgpu_buffer_rsrc_t xr, uint32_t base, uint32_t vs8, uint32_t coff, double x[9]) {
#pragma unroll
  for (int l = 0; l < 9; ++l)
    x[l] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(xr, (int)coff, (int)(base + l * vs8), 0));
}

// acc += B (s dd . sign . x): the block B (LDS) one row at a time, each row's nine FMAs next to its reads
__device__ __forceinline__ void s_link_apply(const double* bk, const double x[9], const double dd[9], double s, uint32_t sgm,
                                             uint32_t chb, double acc[9]) {
  double v[9];
#pragma unroll
  for (int l = 0; l < 9; ++l) {
    const double w = x[l] * (dd[l] * s);
    const int flip = (int)((((sgm >> l) & 1u) ^ chb) << 31);
    v[l] = __hiloint2double(__double2hiint(w) ^ flip, __double2loint(w));
  }
#pragma unroll
  for (int l = 0; l < 9; ++l) {
    double a = acc[l];
#pragma unroll
    for (int l2 = 0; l2 < 9; ++l2) a = fma(bk[l * 10 + l2], v[l2], a);
    acc[l] = a;
    __builtin_amdgcn_sched_barrier(0);
  }
}

template <int NSLOT>
__device__ __forceinline__ void s_links_piid(const OpDev& op, const double* bl, const double* x, uint32_t vs8, uint32_t coff,
                                             int i, uint64_t chim, const double dd[9], bool xnorm, double acc[9]) {
  const uint32_t node8 = 9u * vs8;
  const __amdgpu_buffer_rsrc_t xr =
      __builtin_amdgcn_make_buffer_rsrc((void*)x, (short)0, (int)((uint32_t)(op.nn + 1) * node8), 0x00020000);
  const uint32_t* __restrict__ hp = op.hop + (size_t)i * NSLOT;
  // link k = 0 is the node itself (on-site block), k = 1..NSLOT the bonds; two register sets alternate so that the
  // loads of link k + 1 are in flight during the block product of link k
  double xa[9], xb[9];
  s_link_load<NSLOT>(xr, (uint32_t)i * node8, vs8, coff, xa);
  double sa = xnorm ? op.isq[i] : 1.0, sb = 1.0;
  uint32_t ga = 0u, gb = 0u, ca = 0u, cb = 0u;
#pragma unroll 1
  for (int k = 0; k <= NSLOT; k += 2) {
    if (k + 1 <= NSLOT) {
      const uint32_t e = hp[k];
      const int r = (int)(e >> 6) - 1, g = (int)(e & 63u);
      s_link_load<NSLOT>(xr, (uint32_t)r * node8, vs8, coff, xb);
      sb = xnorm ? op.isq[r] : 1.0;
      gb = op.sgm[g];
      cb = (uint32_t)(chim >> g) & 1u;
    }
    s_link_apply(bl + k * 90, xa, dd, sa, ga, ca, acc);
    if (k + 1 > NSLOT) break;
    if (k + 2 <= NSLOT) {
      const uint32_t e = hp[k + 1];
      const int r = (int)(e >> 6) - 1, g = (int)(e & 63u);
      s_link_load<NSLOT>(xr, (uint32_t)r * node8, vs8, coff, xa);
      sa = xnorm ? op.isq[r] : 1.0;
      ga = op.sgm[g];
      ca = (uint32_t)(chim >> g) & 1u;
    }
    s_link_apply(bl + (k + 1) * 90, xb, dd, sb, gb, cb, acc);
  }
}


template <class T, bool PIID, int NSLOT>
__global__ void __launch_bounds__(BLOCK, 4) k_node(OpDev op, MemDev md, FArgs fa, const double* __restrict__ blk,
                                                   const double* __restrict__ dperm) {
  __shared__ T red[BLOCK];
  extern __shared__ double blds[];
  if (md.tact && !md.tact[blockIdx.y]) return;  // every member of this tile has left
  if constexpr (std::is_same<T, double>::value && PIID && NSLOT > 0) {
    if (fa.xs) stage_blocks(op, blds);
  }
  const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x >> 6;
  const int i = __builtin_amdgcn_readfirstlane(xcd_block(blockIdx.x, gridDim.x) * (BLOCK / WAVE) + wave);
  const int m = blockIdx.y * WAVE + lane;
  const bool mlive = m < md.nb;
  const int nbp = md.nbp;
  T dotv = zero<T>();
  if (i < op.nn) {
    const uint32_t inf = op.info[i];
    const uint64_t chim = md.chimask[m];
    T acc[9];
    for (int l = 0; l < 9; ++l) acc[l] = zero<T>();
    if (fa.xs) {
      const T* __restrict__ x = (const T*)fa.xs;
      const int t = (inf >> 12) & 255u;
      if constexpr (std::is_same<T, double>::value && PIID && NSLOT > 0) {
        double dd[9];
        for (int l = 0; l < 9; ++l) dd[l] = fa.s_din ? md.dvec[(size_t)l * nbp + m] : 1.0;
        s_links_piid<NSLOT>(op, blds + (size_t)t * (NSLOT + 1) * 90, (const double*)x, (uint32_t)nbp * 8u, (uint32_t)m * 8u, i,
                            chim, dd, fa.xs_norm != 0, acc);
      } else {
      const double* __restrict__ b0 = blk + (size_t)t * (op.nslot + 1) * 81;
      {
        const double sc = fa.xs_norm ? op.isq[i] : 1.0;
        const T* __restrict__ row = x + (size_t)i * 9 * nbp;
        T xn[9];
        for (int l = 0; l < 9; ++l) xn[l] = (sc * (fa.s_din ? md.dvec[(size_t)l * nbp + m] : 1.0)) * row[(size_t)l * nbp + m];
        for (int l = 0; l < 9; ++l) {
          T a2 = zero<T>();
          for (int l2 = 0; l2 < 9; ++l2) a2 = fma_s(b0[l * 9 + l2], xn[l2], a2);
          acc[l] = a2;
        }
      }
      const uint32_t* __restrict__ hp = op.hop + (size_t)i * op.nslot;
      const int nslot = NSLOT > 0 ? NSLOT : op.nslot;
#pragma unroll 2
      for (int kk = 0; kk < nslot; ++kk) {
        {
          const uint32_t e = hp[kk];
          const int r = (int)(e >> 6) - 1, g = (int)(e & 63u);
          const double chi = ((chim >> g) & 1ull) ? -1.0 : 1.0;
          const double sc = fa.xs_norm ? op.isq[r] : 1.0;
          const T* __restrict__ row = x + (size_t)r * 9 * nbp;
          T v[9];
          for (int l = 0; l < 9; ++l) {
            const int pl = PIID ? l : op.pi[g * 9 + l];
            const double dd = fa.s_din ? (PIID ? md.dvec[(size_t)l * nbp + m] : dperm[((size_t)g * 9 + l) * nbp + m]) : 1.0;
            v[l] = (chi * op.sg[g * 9 + l] * sc * dd) * row[(size_t)pl * nbp + m];
          }
          const double* __restrict__ bk = b0 + (size_t)(kk + 1) * 81;
          for (int l = 0; l < 9; ++l) {
            T a2 = acc[l];
            for (int l2 = 0; l2 < 9; ++l2) a2 = fma_s(bk[l * 9 + l2], v[l2], a2);
            acc[l] = a2;
          }
        }
      }
      }
      for (int l = 0; l < 9; ++l) acc[l] = (fa.s_coef * (fa.s_dout ? md.dvec[(size_t)l * nbp + m] : 1.0)) * acc[l];
    }
    for (int t = 0; t < fa.nr; ++t)
      rand_accum<T>(op, md, op.pi, op.sg, chim, i, inf, m, fa.qr[t], (const T*)fa.xr[t], fa.r_norm[t], fa.r_coef[t], acc);
    if (fa.nl > 0)
      lin_accum(op, md, op.pi, op.sg, chim, i, inf, m, fa.lq[0], (const T*)fa.xl[0], fa.l_norm[0], fa.l_coef[0], acc);
    if (fa.nl > 1)
      lin_accum(op, md, op.pi, op.sg, chim, i, inf, m, fa.lq[1], (const T*)fa.xl[1], fa.l_norm[1], fa.l_coef[1], acc);
    for (int t = 0; t < fa.no; ++t) {
      const T* xo = (const T*)fa.xo[t];
      for (int l = 0; l < 9; ++l) acc[l] = fma_s(fa.o_coef[t], ld(xo, i, l, nbp, m), acc[l]);
    }
    const double so = fa.out_sq ? op.sq[i] : 1.0;
    T* out = (T*)fa.out;
    const T* du = (const T*)fa.dotu;
    for (int l = 0; l < 9; ++l) {
      T y = so * acc[l];
      if (mlive) {
        out[((size_t)i * 9 + l) * nbp + m] = y;
        if (du) dotv = dotv + bilin(ld(du, i, l, nbp, m), y);
      }
    }
  }
  if (fa.dotu) {
    T* part = (T*)fa.partial;
    block_reduce_member<T>(dotv, WAVE, red, part + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * WAVE + threadIdx.x);
  }
}

// Complex members as pairs of real lanes (lane 2j: Re of member j, lane 2j + 1: Im): everything the real operator does
// acts on each part alone, so the pass runs in FP64 on half the registers of a complex lane; only X(z) = z pa + pb and the
// bilinear dot sum(u w) exchange the partner's part (lane ^ 1).  32 complex members and one node per wave.
__device__ void rand_accum_pair(const OpDev& op, const MemDev& md, uint64_t chim, int i, uint32_t inf, int m, int vi, int vs,
                                int q, const double* __restrict__ x, int xnorm, double coef, double acc[9]) {
  const int nbp = md.nbp, nsm = op.nsmax;
  const int fs = frozen_slot(inf, md, m);
  const double isi = op.isq[i];
  if (fs) {
    const double* fp = frozen_row(md, fs, q);
    for (int l = 0; l < 9; ++l) {
      double xv = x[((size_t)i * 9 + l) * vs + vi];
      if (xnorm) xv = isi * xv;
      acc[l] = fma(coef * fp[(size_t)l * nbp + m], xv, acc[l]);
    }
    return;
  }
  const int l0 = inf & 15u, kap = (inf >> 4) & 255u;
  const double* w = md.wts + (((size_t)q * op.ncls + kap) * nsm + l0) * nsm * 9 * nbp;
  double r_acc[9];
  for (int l = 0; l < 9; ++l) {
    double xv = x[((size_t)i * 9 + l) * vs + vi];
    if (xnorm) xv = isi * xv;
    r_acc[l] = w[((size_t)l0 * 9 + l) * nbp + m] * xv;
  }
  const int nsc = op.nsc[kap];
  const uint32_t* fl = op.flip + (size_t)i * nsm;
  for (int k = 0; k < nsc; ++k) {
    const uint32_t e = fl[k];
    if (k == l0 || e == 0u) continue;
    int r, g;
    unpack_link(e, r, g);
    const double chi = ((chim >> g) & 1ull) ? -1.0 : 1.0;
    const double isr = op.isq[r];
    for (int l = 0; l < 9; ++l) {
      double xv = x[((size_t)r * 9 + op.pi[g * 9 + l]) * vs + vi];
      if (xnorm) xv = isr * xv;
      r_acc[l] = fma(w[((size_t)k * 9 + l) * nbp + m], (chi * op.sg[g * 9 + l]) * xv, r_acc[l]);
    }
  }
  for (int l = 0; l < 9; ++l) acc[l] = fma(coef, r_acc[l], acc[l]);
}

template <bool PIID, int NSLOT>
__global__ void __launch_bounds__(BLOCK, 4) k_pair(OpDev op, MemDev md, FArgs fa, const double* __restrict__ blk,
                                                   const double* __restrict__ dperm) {
  __shared__ cplx red[BLOCK / WAVE][WAVE / 2];
  extern __shared__ double blds[];
  if (md.tact && !md.tact[blockIdx.y / 2]) return;  // every member of this 64-member tile has left
  if constexpr (PIID && NSLOT > 0) {
    if (fa.xs) stage_blocks(op, blds);
  }
  const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x >> 6;
  const int i = __builtin_amdgcn_readfirstlane(xcd_block(blockIdx.x, gridDim.x) * (BLOCK / WAVE) + wave);
  const int part = lane & 1;
  const int m = blockIdx.y * (WAVE / 2) + (lane >> 1);
  const int vs = 2 * md.nbp, vi = 2 * m + part;
  const bool mlive = m < md.nb;
  const int nbp = md.nbp;
  double cre = 0.0, cim = 0.0;
  if (i < op.nn) {
    const uint32_t inf = op.info[i];
    const uint64_t chim = md.chimask[m];
    double acc[9];
    for (int l = 0; l < 9; ++l) acc[l] = 0.0;
    if (fa.xs) {
      const double* __restrict__ x = (const double*)fa.xs;
      const int t = (inf >> 12) & 255u;
      if constexpr (PIID && NSLOT > 0) {
        double dd[9];
        for (int l = 0; l < 9; ++l) dd[l] = fa.s_din ? md.dvec[(size_t)l * nbp + m] : 1.0;
        s_links_piid<NSLOT>(op, blds + (size_t)t * (NSLOT + 1) * 90, x, (uint32_t)vs * 8u, (uint32_t)vi * 8u, i, chim, dd,
                            fa.xs_norm != 0, acc);
      } else {
      const double* __restrict__ b0 = blk + (size_t)t * (op.nslot + 1) * 81;
      {
        const double sc = fa.xs_norm ? op.isq[i] : 1.0;
        const double* __restrict__ row = x + (size_t)i * 9 * vs;
        double xn[9];
        for (int l = 0; l < 9; ++l) xn[l] = (sc * (fa.s_din ? md.dvec[(size_t)l * nbp + m] : 1.0)) * row[(size_t)l * vs + vi];
        for (int l = 0; l < 9; ++l) {
          double a2 = 0.0;
          for (int l2 = 0; l2 < 9; ++l2) a2 = fma(b0[l * 9 + l2], xn[l2], a2);
          acc[l] = a2;
        }
      }
      const uint32_t* __restrict__ hp = op.hop + (size_t)i * op.nslot;
      const int nslot = NSLOT > 0 ? NSLOT : op.nslot;
#pragma unroll 2
      for (int kk = 0; kk < nslot; ++kk) {
        {
          const uint32_t e = hp[kk];
          const int r = (int)(e >> 6) - 1, g = (int)(e & 63u);
          const double chi = ((chim >> g) & 1ull) ? -1.0 : 1.0;
          const double sc = fa.xs_norm ? op.isq[r] : 1.0;
          const double* __restrict__ row = x + (size_t)r * 9 * vs;
          double v[9];
          for (int l = 0; l < 9; ++l) {
            const int pl = PIID ? l : op.pi[g * 9 + l];
            const double dd = fa.s_din ? (PIID ? md.dvec[(size_t)l * nbp + m] : dperm[((size_t)g * 9 + l) * nbp + m]) : 1.0;
            v[l] = (chi * op.sg[g * 9 + l] * sc * dd) * row[(size_t)pl * vs + vi];
          }
          const double* __restrict__ bk = b0 + (size_t)(kk + 1) * 81;
          for (int l = 0; l < 9; ++l) {
            double a2 = acc[l];
            for (int l2 = 0; l2 < 9; ++l2) a2 = fma(bk[l * 9 + l2], v[l2], a2);
            acc[l] = a2;
          }
        }
      }
      }
      for (int l = 0; l < 9; ++l) acc[l] *= fa.s_coef * (fa.s_dout ? md.dvec[(size_t)l * nbp + m] : 1.0);
    }
    for (int t = 0; t < fa.nr; ++t)
      rand_accum_pair(op, md, chim, i, inf, m, vi, vs, fa.qr[t], (const double*)fa.xr[t], fa.r_norm[t], fa.r_coef[t], acc);
    for (int t = 0; t < fa.nl; ++t) {
      double ta[9], tb[9];
      for (int l = 0; l < 9; ++l) { ta[l] = 0.0; tb[l] = 0.0; }
      rand_accum_pair(op, md, chim, i, inf, m, vi, vs, fa.lq[t], (const double*)fa.xl[t], fa.l_norm[t], 1.0, ta);
      rand_accum_pair(op, md, chim, i, inf, m, vi, vs, fa.lq[t] + 1, (const double*)fa.xl[t], fa.l_norm[t], 1.0, tb);
      const cplx z = md.z[m];
      const double sgn = part ? 1.0 : -1.0;
      for (int l = 0; l < 9; ++l) {
        const double oth = __shfl_xor(ta[l], 1);
        acc[l] = fma(fa.l_coef[t], fma(z.re, ta[l], fma(sgn * z.im, oth, tb[l])), acc[l]);
      }
    }
    for (int t = 0; t < fa.no; ++t) {
      const double* xo = (const double*)fa.xo[t];
      for (int l = 0; l < 9; ++l) acc[l] = fma(fa.o_coef[t], xo[((size_t)i * 9 + l) * vs + vi], acc[l]);
    }
    const double so = fa.out_sq ? op.sq[i] : 1.0;
    double* out = (double*)fa.out;
    const double* du = (const double*)fa.dotu;
    for (int l = 0; l < 9; ++l) {
      const double y = so * acc[l];
      if (mlive) out[((size_t)i * 9 + l) * vs + vi] = y;
      if (du) {
        const double yo = __shfl_xor(y, 1);
        const double u = mlive ? du[((size_t)i * 9 + l) * vs + vi] : 0.0;
        const double p1 = u * y;
        cre += part ? -p1 : p1;
        cim = fma(u, yo, cim);
      }
    }
  }
  if (fa.dotu) {
    cre += __shfl_xor(cre, 1);
    cim += __shfl_xor(cim, 1);
    if (!part) red[wave][lane >> 1] = cplx{cre, cim};
    __syncthreads();
    if (threadIdx.x < WAVE / 2) {
      cplx s2 = red[0][threadIdx.x];
      for (int w = 1; w < BLOCK / WAVE; ++w) s2 = s2 + red[w][threadIdx.x];
      ((cplx*)fa.partial)[((size_t)blockIdx.y * gridDim.x + blockIdx.x) * (WAVE / 2) + threadIdx.x] = s2;
    }
  }
}

// ------------------------------------------------------------------------------------------------ adjoint (real chains)
// Node/member mapping of the elementwise adjoint kernels: as k_update (64 / mt nodes per wave, mt members).
#define ASR_NODE_MEMBER                                                                              \
  const int mt = md.mt, npw = WAVE / mt;                                                             \
  const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x >> 6;                                \
  const int m = blockIdx.y * mt + (lane & (mt - 1));                                                 \
  const int i = (xcd_block(blockIdx.x, gridDim.x) * (BLOCK / WAVE) + wave) * npw + lane / mt;        \
  const bool live = (i < op.nn) && (m < md.nb);

// out = sum_t c_t[m] v_t (t < nt, c_t per member, or 1 when null), optional partial dots <out, p_q> (q < np)
__global__ void __launch_bounds__(BLOCK) k_lincomb(OpDev op, MemDev md, double* out, int nt, const double* v0,
                                                   const double* c0, const double* v1, const double* c1, const double* v2,
                                                   const double* c2, const double* v3, const double* c3, int np,
                                                   const double* p0, double* part0, const double* p1, double* part1) {
  __shared__ double red[BLOCK];
  ASR_NODE_MEMBER
  double d0 = 0.0, d1 = 0.0;
  if (live) {
    const double* vv[4] = {v0, v1, v2, v3};
    const double* cc[4] = {c0, c1, c2, c3};
    double co[4];
    for (int t = 0; t < 4; ++t) co[t] = (t < nt) ? (cc[t] ? cc[t][m] : 1.0) : 0.0;
    for (int l = 0; l < 9; ++l) {
      const size_t q = ((size_t)i * 9 + l) * md.nbp + m;
      double y = 0.0;
      for (int t = 0; t < nt; ++t) y = fma(co[t], vv[t][q], y);
      out[q] = y;
      if (np > 0) d0 = fma(p0[q], y, d0);
      if (np > 1) d1 = fma(p1[q], y, d1);
    }
  }
  if (np > 0) block_reduce_member<double>(d0, mt, red, part0 + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * mt + threadIdx.x);
  if (np > 1) {
    __syncthreads();
    block_reduce_member<double>(d1, mt, red, part1 + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * mt + threadIdx.x);
  }
}

// gradient contributions of one level, per member: with x = u_k, y = w_bar = r + abar x (normalised), x~ = x / sq,
// y~ = y / sq, t2x = S D x~, t2y = S D y~ = t2r + abar t2x (amplitude), and R_sl the species-s, orbital-l projection of
// the random operators,
//   dL/dc_sL += sum_i sq_i y_i . (R_sL x~)_i,   dL/dd_sL += sum_i sq_i^2 [(R_sL y~)_i . t2x_i + t2y_i . (R_sL x~)_i]
// (the normalised operators are symmetric on the symmetric subspace).  L = l shell (s, p, d).  R_s gathers the node's own
// state and its flip partners with the class-table weights mt(l0, k, s): every partner is read once and its three
// per-shell products are spread over the species.  A block walks the node blocks blockIdx.x, blockIdx.x + gridDim.x, ...
// of the nbx node blocks (gridDim.x a multiple of the 8 XCDs, so every block stays in its XCD's node range), sums its
// nodes in registers, and adds its 6 NSP sums per member to part (one slot per block and member, planes q = 0..3 NSP-1
// for c, 3 NSP..6 NSP-1 for d) after one reduction for all of them: the sum over levels stays in place until one final
// reduction.
template <int NSP, bool BUF>
__global__ void __launch_bounds__(BLOCK) k_adj_grad(OpDev op, MemDev md, const double* x, const double* r,
                                                    const double* abar, const double* t2x, const double* t2r, const int* spc,
                                                    double* part, int nbx) {
  constexpr int NQ = 6 * NSP, CH = 6;  // quantities per member; reduced CH at a time through LDS
  __shared__ double red[BLOCK / WAVE][CH][WAVE];
  const int mt = md.mt, npw = WAVE / mt;
  const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x >> 6;
  const int m = blockIdx.y * mt + (lane & (mt - 1));
  double g[NQ];  // gc[s][L] at 3 s + L, gd[s][L] at 3 NSP + 3 s + L
#pragma unroll
  for (int q = 0; q < NQ; ++q) g[q] = 0.0;
  const int nbp = md.nbp, nsm = op.nsmax;
  // BUF (vectors below 4 GiB): the own rows as raw buffer loads, one 32-bit offset per lane and the orbital in the scalar
  // offset, so the 36 loads of a node are in flight together at a small register cost
  const int nrec = (int)((uint32_t)(op.nn + 1) * 9u * (uint32_t)nbp * 8u);
  __amdgpu_buffer_rsrc_t rx, rr, rtx, rtr;
  if constexpr (BUF) {
    rx = __builtin_amdgcn_make_buffer_rsrc((void*)x, (short)0, nrec, 0x00020000);
    rr = __builtin_amdgcn_make_buffer_rsrc((void*)r, (short)0, nrec, 0x00020000);
    rtx = __builtin_amdgcn_make_buffer_rsrc((void*)t2x, (short)0, nrec, 0x00020000);
    rtr = __builtin_amdgcn_make_buffer_rsrc((void*)t2r, (short)0, nrec, 0x00020000);
  }
  for (int vb = blockIdx.x; vb < nbx; vb += gridDim.x) {
    const int i = (xcd_block(vb, nbx) * (BLOCK / WAVE) + wave) * npw + lane / mt;
    if (i >= op.nn || m >= md.nb) continue;
    const uint32_t inf = op.info[i];
    const double sqi = op.sq[i], isq = op.isq[i], ab = abar[m];
    const uint64_t chim = md.chimask[m];
    // one pass per gathered term (the node's own state, then every flip partner): px = R-term of x~, py = of y~, and the
    // per-shell products A = y . px, B = t2x . py + t2y . px spread over the species slots with the weights mt(l0, k, s)
    const int fs = frozen_slot(inf, md, m);
    const bool fz = fs != 0;
    const int l0 = inf & 15u, kap = (inf >> 4) & 255u;
    const int nsc = fz ? 1 : op.nsc[kap];
    const double* mtk = md.mtm + ((size_t)kap * nsm + l0) * nsm * nsm * nbp;  // [k][sl][nbp]
    const uint32_t* fl = op.flip + (size_t)i * nsm;
    // the sums of one gathered term into the gradient slots: per species q the weight (0 for the other species), so that
    // every index into g is static and g stays in registers; adding w A with w = 0 leaves a slot unchanged
    auto spread = [&](int k, double A0, double A1, double A2, double B0, double B1, double B2) {
      double wq[NSP];
#pragma unroll
      for (int q = 0; q < NSP; ++q) wq[q] = 0.0;
      if (fz) {
        // a frozen site carries its species: R_s is the identity there for that species alone
        const int s = md.fsp[(size_t)(fs - 1) * nbp + m];
#pragma unroll
        for (int q = 0; q < NSP; ++q) wq[q] = q == s ? 1.0 : 0.0;
#pragma unroll
        for (int q = 0; q < NSP; ++q) {
          g[3 * q + 0] += sqi * wq[q] * A0;
          g[3 * q + 1] += sqi * wq[q] * A1;
          g[3 * q + 2] += sqi * wq[q] * A2;
          g[3 * NSP + 3 * q + 0] += sqi * sqi * wq[q] * B0;
          g[3 * NSP + 3 * q + 1] += sqi * sqi * wq[q] * B1;
          g[3 * NSP + 3 * q + 2] += sqi * sqi * wq[q] * B2;
        }
      } else {
#pragma unroll
        for (int sl = 0; sl < NSP; ++sl) {
          if (sl < op.nsc[kap]) {
            const double w = mtk[((size_t)k * nsm + sl) * nbp + m];
            const int s = spc[kap * nsm + sl];
#pragma unroll
            for (int q = 0; q < NSP; ++q) {
              const double ww = q == s ? w : 0.0;
              g[3 * q + 0] += sqi * ww * A0;
              g[3 * q + 1] += sqi * ww * A1;
              g[3 * q + 2] += sqi * ww * A2;
              g[3 * NSP + 3 * q + 0] += sqi * sqi * ww * B0;
              g[3 * NSP + 3 * q + 1] += sqi * sqi * ww * B1;
              g[3 * NSP + 3 * q + 2] += sqi * sqi * ww * B2;
            }
          }
        }
      }
    };
    // the own state: the gathered values are the node's own x and r; three orbitals' rows in flight at a time
    {
      const uint32_t voff = ((uint32_t)i * 9u * (uint32_t)nbp + (uint32_t)m) * 8u, ls = (uint32_t)nbp * 8u;
      double A[3] = {0.0, 0.0, 0.0}, B[3] = {0.0, 0.0, 0.0};
#pragma unroll 3
      for (int l = 0; l < 9; ++l) {
        double xo, ro, txl, trl;
        if constexpr (BUF) {
          xo = tile_row<double>(rx, voff, l * ls);
          ro = tile_row<double>(rr, voff, l * ls);
          txl = tile_row<double>(rtx, voff, l * ls);
          trl = tile_row<double>(rtr, voff, l * ls);
        } else {
          xo = ld(x, i, l, nbp, m);
          ro = ld(r, i, l, nbp, m);
          txl = ld(t2x, i, l, nbp, m);
          trl = ld(t2r, i, l, nbp, m);
        }
        const int L = l == 0 ? 0 : (l < 4 ? 1 : 2);
        const double yol = ro + ab * xo;
        const double tyl = trl + ab * txl;
        const double pxl = isq * xo;
        const double pyl = isq * (ro + ab * xo);
        const double a0 = fma(yol, pxl, L == 0 ? A[0] : (L == 1 ? A[1] : A[2]));
        const double b0 = fma(txl, pyl, fma(tyl, pxl, L == 0 ? B[0] : (L == 1 ? B[1] : B[2])));
        if (L == 0) { A[0] = a0; B[0] = b0; } else if (L == 1) { A[1] = a0; B[1] = b0; } else { A[2] = a0; B[2] = b0; }
      }
      spread(l0, A[0], A[1], A[2], B[0], B[1], B[2]);
    }
    // the flip partners (the 0.7 % of nodes that have them): the own rows again (cache hits) and the partner's rows
    for (int kk = 0; kk < nsc; ++kk) {
      if (fz || kk == l0) continue;
      const uint32_t e = fl[kk];
      if (e == 0u) continue;
      int rn, gop;
      unpack_link(e, rn, gop);
      const double cs0 = (((chim >> gop) & 1ull) ? -1.0 : 1.0) * op.isq[rn];
      const int* pg = op.pi + gop * 9;
      const double* sgg = op.sg + gop * 9;
      double A[3] = {0.0, 0.0, 0.0}, B[3] = {0.0, 0.0, 0.0};
#pragma unroll 1
      for (int l = 0; l < 9; ++l) {
        const int L = l == 0 ? 0 : (l < 4 ? 1 : 2);
        const double xo = ld(x, i, l, nbp, m), ro = ld(r, i, l, nbp, m);
        const double yol = ro + ab * xo;
        const double txl = ld(t2x, i, l, nbp, m);
        const double tyl = ld(t2r, i, l, nbp, m) + ab * txl;
        const int pl = pg[l];
        const double cs = cs0 * sgg[l];
        const double xv = ld(x, rn, pl, nbp, m);
        const double pxl = cs * xv;
        const double pyl = cs * (ld(r, rn, pl, nbp, m) + ab * xv);
        const double a0 = fma(yol, pxl, L == 0 ? A[0] : (L == 1 ? A[1] : A[2]));
        const double b0 = fma(txl, pyl, fma(tyl, pxl, L == 0 ? B[0] : (L == 1 ? B[1] : B[2])));
        if (L == 0) { A[0] = a0; B[0] = b0; } else if (L == 1) { A[1] = a0; B[1] = b0; } else { A[2] = a0; B[2] = b0; }
      }
      spread(kk, A[0], A[1], A[2], B[0], B[1], B[2]);
    }
  }
  // lanes of one member (lane % mt) summed over the wave's nodes, then over the waves in order, CH quantities per round
  const size_t base = ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * mt;
  const size_t plane = (size_t)gridDim.x * gridDim.y * mt;
#pragma unroll
  for (int q = 0; q < NQ; ++q)
    for (int off = mt; off < WAVE; off <<= 1) g[q] += shfl_xor(g[q], off);
#pragma unroll
  for (int q0 = 0; q0 < NQ; q0 += CH) {
    if (lane < mt) {
#pragma unroll
      for (int c = 0; c < CH; ++c) red[wave][c][lane] = g[q0 + c];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < CH * mt; t += BLOCK) {
      const int c = t / mt, ml = t % mt;
      double s = red[0][c][ml];
      for (int w = 1; w < BLOCK / WAVE; ++w) s += red[w][c][ml];
      double* p = part + (size_t)(q0 + c) * plane + base + ml;
      *p = *p + s;
    }
    __syncthreads();
  }
}

// u_{k+1} = (w - a_k u_k - b_{k-1} u_{k-1}) / b_k of a recomputed level, the arithmetic of k_update (projection off) and
// k_scale element for element, so that a recomputed vector equals the forward's bit for bit
__global__ void __launch_bounds__(BLOCK) k_recompute(OpDev op, MemDev md, double* out, const double* w, const double* u,
                                                     const double* up, const double* a_eff, const double* b_prev,
                                                     const double* inv_b) {
  ASR_NODE_MEMBER
  if (!live) return;
  const double a = a_eff[m], b = b_prev[m], ib = inv_b[m];
  for (int l = 0; l < 9; ++l) {
    const size_t q = ((size_t)i * 9 + l) * md.nbp + m;
    const double xv = w[q] - bilin(a, u[q]) - bilin(b, up[q]);
    out[q] = bilin(ib, xv);
  }
}

// The first pass of H r_bar with r_bar itself: r_bar = alpha u_bar_{k+1} + beta u_{k+1} (k_lincomb's operation order) is
// formed at the rows the pass reads, written once (its own rows), with the partials of <r_bar, u_k> and <r_bar, u_k-1>,
// and t1 = rand_d(r_bar~) (rand_ct's order): the rows and partials of the two separate passes bit for bit, without the
// write and re-read of r_bar between them
__global__ void __launch_bounds__(BLOCK) k_adj_dpass(OpDev op, MemDev md, const double* __restrict__ ub,
                                                     const double* __restrict__ u2, const double* __restrict__ alpha,
                                                     const double* __restrict__ beta, const double* __restrict__ u1,
                                                     const double* __restrict__ u0, double* __restrict__ rb,
                                                     double* __restrict__ t1, double* part2, double* part3) {
  __shared__ double red[BLOCK];
  ASR_NODE_MEMBER
  double d0 = 0.0, d1 = 0.0;
  if (live) {
    const int nbp = md.nbp;
    const double al = alpha[m], be = beta[m];
    auto xrow = [&](int r, int l) -> double {
      const size_t q = ((size_t)r * 9 + l) * nbp + m;
      return fma(be, u2[q], fma(al, ub[q], 0.0));
    };
    double acc[9];
#pragma unroll
    for (int l = 0; l < 9; ++l) {
      const size_t q = ((size_t)i * 9 + l) * nbp + m;
      const double y = xrow(i, l);
      rb[q] = y;
      d0 = fma(u1[q], y, d0);
      d1 = fma(u0[q], y, d1);
      acc[l] = 0.0;
    }
    rand_node_rows<double>(op, md, i, op.info[i], m, xrow, QD, 1.0, 1, false, acc);
#pragma unroll
    for (int l = 0; l < 9; ++l) t1[((size_t)i * 9 + l) * nbp + m] = acc[l];
  }
  block_reduce_member<double>(d0, mt, red, part2 + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * mt + threadIdx.x);
  __syncthreads();
  block_reduce_member<double>(d1, mt, red, part3 + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * mt + threadIdx.x);
}

// per member: the forward coefficients of level k (a_k, b_{k-1}, 1 / b_k, -a_k, -b_k)
__global__ void k_adj_coef(int nb, int ld, int k, const double* a, const double* b, double* a_eff, double* b_prev,
                           double* inv_b, double* nega, double* negb) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nb) return;
  a_eff[m] = a[(size_t)m * ld + k];
  b_prev[m] = k > 0 ? b[(size_t)m * ld + k - 1] : 0.0;
  inv_b[m] = 1.0 / b[(size_t)m * ld + k];
  nega[m] = -a[(size_t)m * ld + k];
  negb[m] = -b[(size_t)m * ld + k];
}

// chain state of a batch
struct ChainDev {
  int* level;        // current level (device scalar, so a captured level step can be replayed)
  int* depth;        // nbp
  int* active;       // nbp
  int* mused;        // nbp
  int ld;            // leading dimension of a, b
  int real_chain;    // 1: real Lanczos conventions (b of the last level recorded)
  double tiny;
  void* cvec;        // route ed3: c_k = <t|u_k> (nbp x ld, the chain's element type) or null
  int* closed;       // route ed3: 1 where b_k < tiny ended the chain (Krylov space exhausted) or null
  const int* s0;     // window runs: the level at which the member of each slot started (its level is level - s0), or null
};

template <class T> __device__ inline double mag(T v);
template <> __device__ inline double mag<double>(double v) { return fabs(v); }
template <> __device__ inline double mag<cplx>(cplx v) { return hypot(v.re, v.im); }
__device__ inline double csqrt_re(double v) { return sqrt(v); }
__device__ inline cplx csqrt_c(cplx v) {
  // principal square root
  double r = hypot(v.re, v.im);
  double a = sqrt(0.5 * (r + fabs(v.re)));
  if (a == 0.0) return {0.0, 0.0};
  if (v.re >= 0.0) return {a, v.im / (2.0 * a)};
  return {fabs(v.im) / (2.0 * a), (v.im >= 0.0 ? a : -a)};
}
__device__ inline double dsqrt(double v) { return csqrt_re(v); }
__device__ inline cplx dsqrt(cplx v) { return csqrt_c(v); }

// fixed-order sum of the per-block partials of member m by one block of BLOCK threads: strided per-thread sums, then a
// shuffle tree per wave and the wave sums in order (deterministic for a given launch shape)
template <class T>
__device__ T block_sum_partials(const T* part, int nbx, int mt, int m, T* lds) {
  const int tile = m / mt, ml = m % mt;
  T s = zero<T>();
  for (int b = threadIdx.x; b < nbx; b += BLOCK) s = s + part[((size_t)tile * nbx + b) * mt + ml];
  for (int off = WAVE / 2; off > 0; off >>= 1) s = s + shfl_xor(s, off);
  const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x >> 6;
  if (lane == 0) lds[wave] = s;
  __syncthreads();
  T tot = lds[0];
  for (int w = 1; w < BLOCK / WAVE; ++w) tot = tot + lds[w];
  return tot;
}

// a_k of every member (one block per member); coefficient used by the update (0 for members that left)
// b_bar_k = dL/db_k + carry - <u_bar_{k+1}, u_{k+1}> / b_k;  r_bar = u_bar_{k+1} / b_k + b_bar_k u_{k+1}
__global__ void __launch_bounds__(BLOCK) k_adj_fin1(int nb, int ld, int k, const double* part, int nbx, int mt,
                                                    const double* b, const double* Lb, const double* cbb, double* alpha,
                                                    double* beta) {
  __shared__ double lds[BLOCK / WAVE];
  const int m = blockIdx.x;
  double d1 = block_sum_partials<double>(part, nbx, mt, m, lds);
  if (threadIdx.x != 0) return;
  const double bk = b[(size_t)m * ld + k];
  alpha[m] = 1.0 / bk;
  beta[m] = Lb[(size_t)m * ld + k] + cbb[m] - d1 / bk;
}

// a_bar_k = dL/da_k - <r_bar, u_k> (and 2 a_bar_k);  the carry to b_bar_{k-1}: -<r_bar, u_{k-1}>
__global__ void __launch_bounds__(BLOCK) k_adj_fin2(int nb, int ld, int k, const double* part2, const double* part3, int nbx,
                                                    int mt, const double* La, double* abar, double* abar2, double* cbb) {
  __shared__ double lds[BLOCK / WAVE];
  const int m = blockIdx.x;
  double d2 = block_sum_partials<double>(part2, nbx, mt, m, lds);
  __syncthreads();
  double d3 = block_sum_partials<double>(part3, nbx, mt, m, lds);
  if (threadIdx.x != 0) return;
  abar[m] = La[(size_t)m * ld + k] - d2;
  abar2[m] = 2.0 * abar[m];
  cbb[m] = -d3;
}

// g[m][q] += sum of the level's partials of gradient component q (fixed order)
__global__ void __launch_bounds__(BLOCK) k_adj_fin_grad(int nb, int nq, const double* part, int nbx, int mt, size_t plane,
                                                        double* g) {
  __shared__ double lds[BLOCK / WAVE];
  const int m = blockIdx.x;
  for (int q = 0; q < nq; ++q) {
    double v = block_sum_partials<double>(part + (size_t)q * plane, nbx, mt, m, lds);
    if (threadIdx.x == 0) g[(size_t)m * nq + q] += v;
    __syncthreads();
  }
}

// sc (lagged normalisation): the chain vector is stored as s_k u_k, so the partials are those of s_k^2 <u, H u>
// lpart (the fused edc level): nlist further partials per member (stride nbp) added after the block partials
// A member that the device depth test of the previous level found converged (conv, set beside that level's update) leaves
// now: the b_k (and route ed3's c_k+1) that the previous level recorded for it are taken back, as if it had left before
// that level's update (the test ran on its own stream, off the level's critical path).
template <class T>
__device__ inline void apply_conv(const ChainDev& cd, int* conv, T* b, int m) {
  conv[m] = 0;
  cd.active[m] = 0;
  const int kp = *cd.level - (cd.s0 ? cd.s0[m] : 0) - 1;
  if (kp >= 0) {
    b[(size_t)m * cd.ld + kp] = zero<T>();
    if (cd.cvec && kp + 1 < cd.ld) ((T*)cd.cvec)[(size_t)m * cd.ld + kp + 1] = zero<T>();
  }
  if (cd.closed) cd.closed[m] = 0;
}

template <class T>
__global__ void k_conv_apply(ChainDev cd, int nb, int* conv, T* b) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < nb && conv[m]) apply_conv<T>(cd, conv, b, m);
}

// Window runs: the members the test of the level just ended found converged leave at once, after that level (the level
// counter already past it), so that their slots take the next chains at the next level instead of running it for
// nothing.  Their slots hold the next vector: flagged like a closed chain's, for a full start column.
template <class T>
__global__ void k_conv_leave(ChainDev cd, int nb, int* conv, T* b) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < nb && conv[m]) {
    apply_conv<T>(cd, conv, b, m);
    cd.closed[m] = 1;
  }
}

template <class T>
__global__ void __launch_bounds__(BLOCK) k_fin_a(ChainDev cd, int nb, const T* part, int nbx, int mt, T* a, T* a_eff,
                                                  const T* sc, const T* lpart, int nlist, int nbp, int* conv, T* bk) {
  __shared__ T lds[BLOCK / WAVE];
  const int m = blockIdx.x;
  int act = cd.active[m];
  if (conv) {
    const int cv = conv[m];
    __syncthreads();  // every thread has read the flag before it is cleared
    if (cv) {
      if (threadIdx.x == 0) apply_conv<T>(cd, conv, bk, m);
      act = 0;
    }
  }
  if (!act) {  // a member that has left: its partials are not read (its tile may not have run)
    if (threadIdx.x == 0) a_eff[m] = zero<T>();
    return;
  }
  T s = block_sum_partials(part, nbx, mt, m, lds);
  if (lpart) {
    __syncthreads();
    T sl = zero<T>();
    for (int j = threadIdx.x; j < nlist; j += BLOCK) sl = sl + lpart[(size_t)j * nbp + m];
    for (int off = WAVE / 2; off > 0; off >>= 1) sl = sl + shfl_xor(sl, off);
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x >> 6;
    if (lane == 0) lds[wave] = sl;
    __syncthreads();
    T tot = lds[0];
    for (int w = 1; w < BLOCK / WAVE; ++w) tot = tot + lds[w];
    s = s + tot;
  }
  if (threadIdx.x != 0) return;
  if (sc) s = bilin(s, inv(bilin(sc[m], sc[m])));
  const int k = *cd.level - (cd.s0 ? cd.s0[m] : 0);
  a[(size_t)m * cd.ld + k] = s;
  cd.mused[m] = k + 1;
  a_eff[m] = s;
  // a complex chain (no projection c_k) at its last level leaves now: its b_k is not recorded, so the update gives it a
  // zero vector (the slot's vector is clean for the next chain of a window run)
  if (!cd.real_chain && !cd.cvec && k == cd.depth[m] - 1) cd.active[m] = 0;
}

// b_k = sqrt(sum w w); members leave at their depth or when b_k < tiny; next u = w * inv_b.  part2 (route ed3): the partials
// of <t, w>, so that c_{k+1} = <t, u_{k+1}> = <t, w> / b_k.  Lagged normalisation (sc non-null): w holds s_k w_k, so
// b_k = sqrt(sum / s_k^2) (the principal root of the same number as the unscaled chain) and the next vector is w itself
// with the scale s_{k+1} = s_k b_k (snext; 1 for a member that left, whose vector the update zeroes).
// pchk (the fused edc level): the partials of the direct <u, H u>; a member whose a_k (taken from the first pass on the
// symmetry of the operator) differs from it beyond rounding sets *asym (the operator is not symmetric on the subspace)
template <class T>
__global__ void __launch_bounds__(BLOCK) k_fin_b(ChainDev cd, int nb, const T* part, int nbx, int mt, T* b, T* b_eff,
                                                  T* inv_b, const T* part2, const T* sc, T* snext, const T* pchk, int* asym,
                                                  const T* a) {
  __shared__ T lds[BLOCK / WAVE];
  const int m = blockIdx.x;
  if (!cd.active[m]) {  // a member that has left: no partials read, a zero update and a unit scale
    if (threadIdx.x == 0) {
      b_eff[m] = zero<T>();
      if (inv_b) inv_b[m] = zero<T>();
      if (snext) {
        T one = zero<T>();
        *(double*)&one = 1.0;
        snext[m] = one;
      }
    }
    return;
  }
  T s = block_sum_partials(part, nbx, mt, m, lds);
  T s2 = zero<T>(), sd = zero<T>();
  if (part2) {
    __syncthreads();
    s2 = block_sum_partials(part2, nbx, mt, m, lds);
  }
  if (pchk) {
    __syncthreads();
    sd = block_sum_partials(pchk, nbx, mt, m, lds);
  }
  if (threadIdx.x != 0) return;
  const int k = *cd.level - (cd.s0 ? cd.s0[m] : 0);
  T s_m = zero<T>();
  if (sc) {
    s_m = sc[m];
    const T is2 = inv(bilin(s_m, s_m));
    s = bilin(s, is2);
    s2 = bilin(s2, inv(s_m));
    sd = bilin(sd, is2);
  }
  if (pchk && cd.active[m]) {
    const T af = a[(size_t)m * cd.ld + k];
    if (!(mag(sd - af) <= 1.0e-9 * mag(sd) + 1.0e-12)) *asym = 1;
  }
  T bk = dsqrt(s);
  int act = cd.active[m];
  if (act) {
    bool last = (k == cd.depth[m] - 1);
    if (!last || cd.real_chain) b[(size_t)m * cd.ld + k] = bk;
    if (!last && mag(bk) < cd.tiny && cd.closed) cd.closed[m] = 1;
    if (last || mag(bk) < cd.tiny) act = 0;
  }
  cd.active[m] = act;
  b_eff[m] = act ? bk : zero<T>();
  if (inv_b) inv_b[m] = act ? inv(bk) : zero<T>();
  if (snext) {
    T one = zero<T>();
    *(double*)&one = 1.0;
    snext[m] = act ? bilin(s_m, bk) : one;
  }
  if (act && part2 && cd.cvec) ((T*)cd.cvec)[(size_t)m * cd.ld + k + 1] = bilin(s2, inv(bk));
}

// ---------------------------------------------------------------------------------------------- adaptive depth on the device
// Eigenvalues of the complex symmetric tridiagonal chain matrix (diagonal d[0..n-1], off-diagonal e[0..n-2], elements
// st apart, both overwritten; the eigenvalues come back in d): the implicit QL iteration with shifts (the method of
// EISPACK tql1) in complex arithmetic with complex orthogonal rotations (c^2 + s^2 = 1), which keep a complex symmetric
// tridiagonal matrix tridiagonal.  Returns 0, or 1 when a rotation breaks down (c^2 + s^2 = 0) or the iteration does not
// converge.
__device__ inline cplx csq(cplx v) { return csqrt_c(v); }
__device__ int cs_tridiag_eigs(int n, cplx* dv, cplx* ev, int st) {
#define d(x) dv[(x) * st]
#define e(x) ev[(x) * st]
  const double eps = 2.220446049250313e-16;
  e(n - 1) = zero<cplx>();
  for (int l = 0; l < n; ++l) {
    int iter = 0, m;
    for (;;) {
      for (m = l; m < n - 1; ++m)
        if (mag(e(m)) <= eps * (mag(d(m)) + mag(d(m + 1)))) break;
      if (m == l) break;
      if (iter++ == 60) return 1;
      cplx g = (d(l + 1) - d(l)) * inv(2.0 * e(l));
      cplx r = csq(g * g + cplx{1.0, 0.0});
      // the root that makes |g + r| large (no cancellation)
      const cplx gp = g + r, gm = g - r;
      g = d(m) - d(l) + e(l) * inv(mag(gp) >= mag(gm) ? gp : gm);
      cplx sn = {1.0, 0.0}, c = {1.0, 0.0}, p = zero<cplx>();
      bool split = false;
      for (int i = m - 1; i >= l; --i) {
        const cplx f = sn * e(i), b = c * e(i);
        r = csq(f * f + g * g);
        e(i + 1) = r;
        if (mag(r) == 0.0) {
          if (mag(f) + mag(g) != 0.0) return 1;  // an isotropic pair: the rotation does not exist
          d(i + 1) = d(i + 1) - p;
          e(m) = zero<cplx>();
          split = true;
          break;
        }
        const cplx ir = inv(r);
        sn = f * ir;
        c = g * ir;
        g = d(i + 1) - p;
        r = (d(i) - g) * sn + 2.0 * (c * b);
        p = sn * r;
        d(i + 1) = g + p;
        g = c * r - b;
      }
      if (split) continue;
      d(l) = d(l) - p;
      e(l) = g;
      e(m) = zero<cplx>();
    }
  }
  return 0;
}
#undef d
#undef e

constexpr int DEPTH_MAX = 128;  // chains of the device depth test: at most this many levels

// G_11 (and f = sum_k c_k G_k1) of a chain of m levels terminated with the tail of cf_species (src/asr_chain.f90:67-167):
// tail 0 hull (the extreme real parts of the chain matrix's eigenvalues give a_inf, b_inf), 2 avg2 (the last two levels),
// 3 none.  Returns 1 when the hull's eigenvalues fail (the caller then keeps the chain running).
// wd, we: work arrays of DEPTH_MAX values at stride st (LDS rows of the calling thread)
__device__ int cf_species_dev(const cplx* a, const cplx* b, const cplx* c, int m, cplx z, int tail, cplx* g11, cplx* f,
                              cplx* wd, cplx* we, int st) {
  cplx aa = a[m - 1], bb = m > 1 ? b[m - 2] : zero<cplx>();
  if (tail == 0 && m >= 2) {
    for (int j = 0; j < m; ++j) { wd[j * st] = a[j]; we[j * st] = j < m - 1 ? b[j] : zero<cplx>(); }
    if (cs_tridiag_eigs(m, wd, we, st)) return 1;
    int lo = 0, hi = 0;
    for (int j = 1; j < m; ++j) {
      if (wd[j * st].re < wd[lo * st].re) lo = j;
      if (wd[j * st].re > wd[hi * st].re) hi = j;
    }
    aa = 0.5 * (wd[lo * st] + wd[hi * st]);
    bb = 0.25 * (wd[hi * st] - wd[lo * st]);
  }
  if (tail == 2 && m >= 3) {
    aa = 0.5 * (a[m - 1] + a[m - 2]);
    bb = csq(0.5 * (b[m - 2] * b[m - 2] + b[m - 3] * b[m - 3]));
  }
  if (tail == 3) { aa = zero<cplx>(); bb = zero<cplx>(); }
  const cplx dz = z - aa;
  cplx tl = zero<cplx>();
  if (mag(bb) >= 1e-14) {
    const cplx s = csq(dz * dz - 4.0 * (bb * bb));
    const cplx ib = inv(2.0 * (bb * bb));
    const cplx g1 = (dz - s) * ib, g2 = (dz + s) * ib;
    tl = mag(bb * g1) <= mag(bb * g2) ? (bb * bb) * g1 : (bb * bb) * g2;
  }
  // G_k down the chain (in wd), then f with the first column pv_k = prod G
  cplx* gt = wd;
  gt[(m - 1) * st] = inv(z - a[m - 1] - tl);
  for (int j = m - 2; j >= 0; --j) gt[j * st] = inv(z - a[j] - (b[j] * b[j]) * gt[(j + 1) * st]);
  cplx prod = gt[0], ff = zero<cplx>();
  for (int k = 0; k < m; ++k) {
    if (c) ff = ff + c[k] * prod;
    if (k < m - 1) prod = prod * b[k] * gt[(k + 1) * st];
  }
  *g11 = gt[0];
  *f = ff;
  return 0;
}

// The adaptive depth of edc_green (src/asr_edc.f90:121-127) at level k1 (1-based; k1 = 8, 12, ...) on the device, one
// thread per member: a member still running below its depth gets G_k1 of its chain (and route ed3's projection f, with
// the check of ed3's chain, src/asr_ed3.f90:177-184), and leaves when |G - G_prev| <= tol |G| (k1 > 8).  gprev / fprev keep
// the previous tested level's values.  It runs on a stream of its own after the a_k reduction, beside the level's update;
// a converged member is flagged (conv) and leaves at the next level's k_fin_a (in a window run at the end of the level,
// k_conv_leave), which takes back what the level recorded for it after a_k: the coefficients and depths are those of the
// test run before the update.
constexpr int DT_TPB = 16;  // threads (members) per block of the device depth test: 2 x 128 x 16 values of LDS
__global__ void __launch_bounds__(DT_TPB) k_depth_test(ChainDev cd, int nb, const cplx* a, const cplx* b, const cplx* cvec,
                                                       const cplx* z, int k1, double tol, int tail, cplx* gprev, cplx* fprev,
                                                       const double* tnorm, int* conv) {
  // the QL iteration's work rows in LDS (a run-time indexed private array would live in scratch memory)
  __shared__ cplx wsd[DEPTH_MAX * DT_TPB], wse[DEPTH_MAX * DT_TPB];
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nb) return;
  // window runs: the member's own level count (the host launches the test at the levels where some member is due)
  if (cd.s0) {
    k1 -= cd.s0[q];
    if (k1 < 8 || (k1 % 4) != 0) return;
  }
  // the members that took part in this level's a_k (mused, written by k_fin_a, is not touched by the level's k_fin_b,
  // which may run at the same time)
  if (cd.mused[q] != k1 || k1 >= cd.depth[q] || k1 > DEPTH_MAX) return;
  cplx g, f;
  if (cf_species_dev(a + (size_t)q * cd.ld, b + (size_t)q * cd.ld, cvec ? cvec + (size_t)q * cd.ld : nullptr, k1, z[q], tail,
                     &g, &f, wsd + threadIdx.x, wse + threadIdx.x, DT_TPB))
    return;
  if (k1 > 8) {
    const double ag = mag(g);
    bool cv = mag(g - gprev[q]) <= tol * ag;
    if (cvec) cv = cv && mag(f - fprev[q]) <= tol * fmax(mag(f), tnorm[q] * ag);
    if (cv) conv[q] = 1;
  }
  gprev[q] = g;
  fprev[q] = f;
}

// lagged normalisation: w <- w / s (the vector of the next level normalised), then s <- 1 (k_scale_one)
template <class T>
__global__ void k_rescale(T* w, const T* s, size_t n, int nbp) {
  size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int m = (int)(q % nbp);
  w[q] = bilin(inv(s[m]), w[q]);
}
template <class T>
__global__ void k_scale_one(T* s, int nbp) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nbp) return;
  T one = zero<T>();
  *(double*)&one = 1.0;
  s[m] = one;
}

// per-member sum of block partials (one block per member)
template <class T>
__global__ void __launch_bounds__(BLOCK) k_sum_member(const T* part, int nbx, int mt, T* out) {
  __shared__ T lds[BLOCK / WAVE];
  const int m = blockIdx.x;
  T s = block_sum_partials(part, nbx, mt, m, lds);
  if (threadIdx.x == 0) out[m] = s;
}

// Start vectors of the two-chain blocks (ed3_start, src/asr_ed3.f90:131-144; epi_start, src/asr_epi.f90:578-587) from
// vv = h'|s0> (projected), s0 the member's start vector on the start node (s0v: 9 x nbp, the unit vector of the start
// orbital or a symmetry-adapted combination): u = su s0 + sv vv, t = tsu s0 + tsv vv, coef = (su, sv, tsu, tsv) per member.
__global__ void k_block_start(const cplx* vv, cplx* u, cplx* tv, const double* coef, const double* s0v, size_t n, int nbp) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int m = (int)(q % nbp);
  const size_t row = q / nbp;  // node * 9 + l: rows 0..8 are the start node
  const double s0 = row < 9 ? s0v[row * nbp + m] : 0.0;
  const double* cf = coef + 4 * (size_t)m;
  const cplx v = vv[q];
  cplx uq = cf[1] * v, tq = cf[3] * v;
  uq.re += cf[0] * s0;
  tq.re += cf[2] * s0;
  u[q] = uq;
  tv[q] = tq;
}

// end of a level: level += 1 and the activity of every member tile (a tile whose members have all left is skipped)
// Member compaction: the permutation map (new slot -> old slot) applied along the member axis of a member-fastest array
// (nrows rows of nbp values), dst[r][q] = src[r][map[q]] for the first ncol slots q of every row, and along the rows of a
// member-major array (nbp rows of ld values), dst[q][k] = src[map[q]][k].
template <class E>
__global__ void k_perm_cols(E* __restrict__ dst, const E* __restrict__ src, size_t nrows, int nbp, int ncol,
                            const int* __restrict__ map) {
  const size_t n = nrows * (size_t)ncol;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    const size_t r = t / (size_t)ncol;
    const int q = (int)(t - r * (size_t)ncol);
    dst[r * (size_t)nbp + q] = src[r * (size_t)nbp + map[q]];
  }
}
template <class E>
__global__ void k_perm_rows(E* __restrict__ dst, const E* __restrict__ src, int ld, int nbp, const int* __restrict__ map) {
  const size_t n = (size_t)nbp * ld;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
    const int q = (int)(t / (size_t)ld), k = (int)(t % (size_t)ld);
    dst[t] = src[(size_t)map[q] * ld + k];
  }
}

// Window runs: the chains wait in a queue and take the slots of members that have left.  k_extract copies the results of
// leaving members (their relative levels 0 .. ld - 1) to the rows of their chains; k_fill_cols copies the member tables
// of the starting chains from the all-chain tables (stride nbps) into their slots (stride nbpd); k_start_cols writes their
// start vectors (u = the unit vector of the start orbital on the start node, rows 0..8; u_prev keeps what the slot held,
// finite values that the first level multiplies by b_-1 = 0); k_member_start resets their recursion state.
template <class T>
__global__ void k_extract(T* __restrict__ a_all, T* __restrict__ b_all, int* __restrict__ mused_all, const T* __restrict__ a,
                          const T* __restrict__ b, const int* __restrict__ mused, int ld, const int* __restrict__ slots,
                          const int* __restrict__ chains, int n, const int* __restrict__ nd = nullptr) {
  if (nd) n = *nd;  // the count from the device queue
  const int j = blockIdx.x;
  if (j >= n) return;
  const int sl = slots[j], c = chains[j];
  for (int k = threadIdx.x; k < ld; k += blockDim.x) {
    a_all[(size_t)c * ld + k] = a[(size_t)sl * ld + k];
    b_all[(size_t)c * ld + k] = b[(size_t)sl * ld + k];
  }
  if (threadIdx.x == 0) mused_all[c] = mused[sl];
}
template <class E>
__global__ void k_fill_cols(E* __restrict__ dst, const E* __restrict__ src, size_t nrows, int nbpd, int nbps,
                            const int* __restrict__ slots, const int* __restrict__ chains, int n,
                            const int* __restrict__ nd = nullptr) {
  if (nd) n = *nd;
  const size_t tot = nrows * (size_t)n;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += (size_t)gridDim.x * blockDim.x) {
    const size_t r = t / (size_t)n;
    const int j = (int)(t - r * (size_t)n);
    dst[r * (size_t)nbpd + slots[j]] = src[r * (size_t)nbps + chains[j]];
  }
}
template <class T>
__global__ void k_start_cols(T* __restrict__ u, size_t nrows, int nbp, const int* __restrict__ slots,
                             const int* __restrict__ l0, int n, const int* __restrict__ nd = nullptr) {
  if (nd) n = *nd;
  const size_t tot = nrows * (size_t)n;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += (size_t)gridDim.x * blockDim.x) {
    const size_t r = t / (size_t)n;
    const int j = (int)(t - r * (size_t)n);
    T v = zero<T>();
    if (r == (size_t)l0[j]) *(double*)&v = 1.0;
    u[r * (size_t)nbp + slots[j]] = v;
  }
}
template <class T>
__global__ void k_member_start(ChainDev cd, const int* __restrict__ slots, const int* __restrict__ depth, int n, int s0,
                               T* a, T* b, T* a_eff, T* b_eff, T* inv_b, T* sc0, T* sc1, T* sc2, cplx* gprev, cplx* fprev,
                               int* s0v, int* closed, int* conv, const int* __restrict__ nd = nullptr) {
  if (nd) {
    // the device queue: the count from the queue, the start level from the level counter
    n = *nd;
    s0 = *cd.level;
  }
  const int j = blockIdx.x;
  if (j >= n) return;
  const int sl = slots[j];
  for (int k = threadIdx.x; k < cd.ld; k += blockDim.x) {
    a[(size_t)sl * cd.ld + k] = zero<T>();
    b[(size_t)sl * cd.ld + k] = zero<T>();
  }
  if (threadIdx.x != 0) return;
  T one = zero<T>();
  *(double*)&one = 1.0;
  a_eff[sl] = zero<T>();
  b_eff[sl] = zero<T>();
  inv_b[sl] = zero<T>();
  sc0[sl] = one;
  sc1[sl] = one;
  sc2[sl] = one;
  if (gprev) gprev[sl] = zero<cplx>();
  if (fprev) fprev[sl] = zero<cplx>();
  cd.depth[sl] = depth[j];
  cd.active[sl] = depth[j] > 0 ? 1 : 0;
  cd.mused[sl] = 0;
  s0v[sl] = s0;
  if (closed) closed[sl] = 0;
  if (conv) conv[sl] = 0;
}

// the start vectors of new members in slots whose vectors are zero (the member before left by the depth test or at its
// depth, so its last update wrote zeros): u = 1 on the start orbital's row of the start node
template <class T>
__global__ void k_start_unit(T* __restrict__ u, int nbp, const int* __restrict__ slots, const int* __restrict__ l0, int n,
                             const int* __restrict__ rng = nullptr) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (rng) {
    // the device queue: entries rng[0] .. rng[1] - 1 of the start list
    j += rng[0];
    n = rng[1];
  }
  if (j >= n) return;
  T v = zero<T>();
  *(double*)&v = 1.0;
  u[(size_t)l0[j] * nbp + slots[j]] = v;
}

// exclusive prefix sum over the threads of a block (in thread order) and the block's total; every thread calls it
constexpr int RF_TPB = 256;
__device__ inline int block_excl_scan(int v, int* wsum, int* total) {
  const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x >> 6;
  int x = v;
  for (int off = 1; off < WAVE; off <<= 1) {
    const int y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  if (lane == WAVE - 1) wsum[wave] = x;
  __syncthreads();
  int before = 0, tot = 0;
  for (int w = 0; w < RF_TPB / WAVE; ++w) {
    if (w < wave) before += wsum[w];
    tot += wsum[w];
  }
  __syncthreads();
  *total = tot;
  return before + x - v;
}

// The window run's queue on the device (one block; the slots in chunks of RF_TPB, slot order kept by prefix sums):
// after a level, the members that have left hand over their slots (extraction list xs -> xc; a slot whose chain ended
// early, closed, keeps a non-zero vector and is flagged dirty), the free slots take the next chains of the queue in slot
// order, and the start list is written dirty slots first (they need a full start column) then clean ones, each in slot
// order: the lists, counts and tile activity of the host's refill, without a host round trip.  wq: [0] next chain,
// [1] extractions, [2] dirty starts, [3] starts, [4] live members, [5] 1 while anything remains to run.  Dynamic LDS:
// nb ints (the slots' activity after the hand-over).
__global__ void __launch_bounds__(RF_TPB) k_refill(ChainDev cd, int nb, int mt, int nchain, int* wq, int* slot_chain,
                                                     int* slot_dirty, const int* nl_all, const int* so_all, int* xs, int* xc,
                                                     int* ws, int* wc, int* wv, int* wl0, int* tact, int ntile) {
  extern __shared__ int live[];
  __shared__ int wsum[RF_TPB / WAVE];
  const int qn = wq[0];
  // the members that have left, in slot order
  int nx = 0;
  for (int b = 0; b < nb; b += RF_TPB) {
    const int q = b + (int)threadIdx.x;
    const int c = q < nb ? slot_chain[q] : -1;
    const int out = (c >= 0 && !cd.active[q]) ? 1 : 0;
    int tot;
    const int p = block_excl_scan(out, wsum, &tot);
    if (out) {
      xs[nx + p] = q;
      xc[nx + p] = c;
      slot_chain[q] = -1;
      if (cd.closed && cd.closed[q]) slot_dirty[q] = 1;
    }
    nx += tot;
  }
  __syncthreads();
  // the free slots, and the number that take a chain now
  int nfree = 0;
  for (int b = 0; b < nb; b += RF_TPB) {
    const int q = b + (int)threadIdx.x;
    int tot;
    (void)block_excl_scan(q < nb && slot_chain[q] < 0 ? 1 : 0, wsum, &tot);
    nfree += tot;
  }
  const int take = min(nfree, max(nchain - qn, 0));
  // the dirty ones among the takers (their starts come first)
  int ndirty = 0;
  {
    int rf = 0;
    for (int b = 0; b < nb; b += RF_TPB) {
      const int q = b + (int)threadIdx.x;
      const int fr = (q < nb && slot_chain[q] < 0) ? 1 : 0;
      int tot, totd;
      const int r = rf + block_excl_scan(fr, wsum, &tot);
      const int dt = (fr && r < take && slot_dirty[q]) ? 1 : 0;
      (void)block_excl_scan(dt, wsum, &totd);
      rf += tot;
      ndirty += totd;
    }
  }
  // the takers in slot order: chain qn + (free rank), start list position by dirtiness
  int rf = 0, rd = 0;
  for (int b = 0; b < nb; b += RF_TPB) {
    const int q = b + (int)threadIdx.x;
    const int fr = (q < nb && slot_chain[q] < 0) ? 1 : 0;
    int tot, totd;
    const int r = rf + block_excl_scan(fr, wsum, &tot);
    const bool taker = fr && r < take;
    const int dirty = (taker && slot_dirty[q]) ? 1 : 0;
    const int rdq = rd + block_excl_scan(dirty, wsum, &totd);
    int lv = 0;
    if (q < nb) lv = (slot_chain[q] >= 0 && cd.active[q]) ? 1 : 0;
    if (taker) {
      const int c = qn + r;
      const int j = dirty ? rdq : ndirty + (r - rdq);
      ws[j] = q;
      wc[j] = c;
      wv[j] = nl_all[c];
      wl0[j] = so_all[c];
      lv = nl_all[c] > 0 ? 1 : 0;
    }
    __syncthreads();  // every thread has read slot_chain and slot_dirty of this chunk
    if (taker) {
      slot_chain[q] = qn + r;
      slot_dirty[q] = 0;
    }
    if (q < nb) live[q] = lv;
    rf += tot;
    rd += totd;
  }
  __syncthreads();
  // running members per tile and in all
  int nl = 0;
  for (int b = 0; b < nb; b += RF_TPB) {
    const int q = b + (int)threadIdx.x;
    int tot;
    (void)block_excl_scan(q < nb ? live[q] : 0, wsum, &tot);
    nl += tot;
  }
  for (int t = (int)threadIdx.x; t < ntile; t += RF_TPB) {
    int any = 0;
    for (int q = t * mt; q < min(nb, (t + 1) * mt); ++q) any |= live[q];
    tact[t] = any;
  }
  if (threadIdx.x == 0) {
    wq[0] = qn + take;
    wq[1] = nx;
    wq[2] = ndirty;
    wq[3] = take;
    wq[4] = nl;
    wq[5] = (nl > 0 || qn + take < nchain) ? 1 : 0;
  }
}

__global__ void k_level_end(int* level, const int* active, int nb, int mt, int* tact, int ntile) {
  for (int t = threadIdx.x; t < ntile; t += blockDim.x) {
    int any = 0;
    for (int q = t * mt; q < std::min(nb, (t + 1) * mt); ++q) any |= active[q];
    tact[t] = any;
  }
  if (threadIdx.x == 0) *level += 1;
}

// w = w - a u - b_prev up; optional projection onto the symmetric subspace; partial sums w w and (tv non-null) tv w
// NT: streaming loads and stores (the member tile fills whole 128-byte lines per row segment; at smaller tiles the partial
// lines are better left to the caches)
template <class T, bool NT>
__global__ void __launch_bounds__(BLOCK) k_update(OpDev op, MemDev md, T* w, const T* u, const T* up, const T* a_eff,
                                                   const T* b_prev, int project, T* partial, const T* tv, T* partial2,
                                                   const T* sc, const T* scp, const int* active) {
  __shared__ T red[BLOCK];
  if (md.tact && !md.tact[blockIdx.y]) return;  // every member of this tile has left
  const int mt = md.mt, npw = WAVE / mt;
  const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x >> 6;
  const int m = blockIdx.y * mt + (lane & (mt - 1));
  const int i = (xcd_block(blockIdx.x, gridDim.x) * (BLOCK / WAVE) + wave) * npw + lane / mt;
  T dv = zero<T>(), dv2 = zero<T>();
  if (i < op.nn && m < md.nb) {
    // lagged normalisation (sc, scp: scales of u and up): w - a u - b_prev (s_k / s_{k-1}) up keeps the scale s_k; a member
    // that has left gets a zero vector
    // (complex chains only: the real chains keep k_scale and their register budget)
    const T a = a_eff[m];
    T b = b_prev[m];
    bool off = false;
    if constexpr (std::is_same<T, cplx>::value) {
      if (sc) b = bilin(b, bilin(sc[m], inv(scp[m])));
      off = active && !active[m];
    }
    T x[9];
    for (int l = 0; l < 9; ++l) {
      size_t q = ((size_t)i * 9 + l) * md.nbp + m;
      if constexpr (NT) x[l] = off ? zero<T>() : nt_load(&w[q]) - bilin(a, nt_load(&u[q])) - bilin(b, nt_load(&up[q]));
      else x[l] = off ? zero<T>() : w[q] - bilin(a, u[q]) - bilin(b, up[q]);
    }
    if (project) {
      int s0 = op.stab_ptr[i], s1 = op.stab_ptr[i + 1], ns = s1 - s0;
      if (std::is_same<T, cplx>::value && ns > 1 && op.piid) {
        // every operation keeps each orbital: (P x)_l = x_l sum_g chi(g) sg(g, l) / |S| (no re-read of the row)
        const uint64_t chim = md.chimask[m];
        double f[9];
        for (int l = 0; l < 9; ++l) f[l] = 0.0;
        for (int j = s0; j < s1; ++j) {
          const int g = op.stab[j];
          const double chi = ((chim >> g) & 1ull) ? -1.0 : 1.0;
          for (int l = 0; l < 9; ++l) f[l] += chi * op.sg[g * 9 + l];
        }
        for (int l = 0; l < 9; ++l) x[l] = (f[l] / (double)ns) * x[l];
      } else if (ns > 1) {
        const uint64_t chim = md.chimask[m];
        T acc[9];
        for (int l = 0; l < 9; ++l) acc[l] = zero<T>();
        for (int j = s0; j < s1; ++j) {
          int g = op.stab[j];
          double chi = ((chim >> g) & 1ull) ? -1.0 : 1.0;
          for (int l = 0; l < 9; ++l) {
            // x[pi(l)] recomputed from memory (w is written below): a register array indexed at run time would live
            // in scratch memory
            const size_t q2 = ((size_t)i * 9 + op.pi[g * 9 + l]) * md.nbp + m;
            const T xv = off ? zero<T>() : w[q2] - bilin(a, u[q2]) - bilin(b, up[q2]);
            acc[l] = acc[l] + (chi * op.sg[g * 9 + l]) * xv;
          }
        }
        for (int l = 0; l < 9; ++l) x[l] = acc[l] / (double)ns;
      }
    }
    // the dots (route ed3's tv rows read) before the first store: a load behind a store waits for it as well
    for (int l = 0; l < 9; ++l) {
      dv = dv + bilin(x[l], x[l]);
      if constexpr (std::is_same<T, cplx>::value) {
        if (tv) dv2 = dv2 + bilin(tv[((size_t)i * 9 + l) * md.nbp + m], x[l]);
      }
    }
    for (int l = 0; l < 9; ++l) {
      const size_t q = ((size_t)i * 9 + l) * md.nbp + m;
      if constexpr (NT) nt_store(&w[q], x[l]);  // the next level gathers it: no use in L2 now
      else w[q] = x[l];
    }
  }
  block_reduce_member<T>(dv, mt, red, partial + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * mt + threadIdx.x);
  if constexpr (std::is_same<T, cplx>::value) {
    if (tv) {
      __syncthreads();
      block_reduce_member<T>(dv2, mt, red, partial2 + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * mt + threadIdx.x);
    }
  }
}

// Projections of the chain vector on the target states (edc_green_proj, src/asr_edc.f90:197-204): for target t and
// orbital l, cpr(t, l) = sum_ks tw(t, ks, l) chi(g) sg(l, g) u(r, pi(l, g)) / sq(r) with (r, g) = (tnode, tgop)(t, ks)
// (augcls_amp_c).  One thread per (member, t, l); members fastest, so the reads of a node row coalesce.
template <class T>
__global__ void k_proj(OpDev op, MemDev md, const T* __restrict__ u, int nt, const int* __restrict__ tnode,
                       const int* __restrict__ tgop, const double* __restrict__ tw, T* __restrict__ cpr,
                       const T* __restrict__ sc) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  const int tl = blockIdx.y, t = tl / 9, l = tl % 9;
  if (m >= md.nb) return;
  const int nsm = op.nsmax, nbp = md.nbp;
  const uint64_t chim = md.chimask[m];
  T acc = zero<T>();
  for (int ks = 0; ks < nsm; ++ks) {
    const int r = tnode[t * nsm + ks];
    if (r < 0) continue;
    const int g = tgop[t * nsm + ks];
    const double cs = (((chim >> g) & 1ull) ? -1.0 : 1.0) * op.sg[g * 9 + l] * op.isq[r];
    const double w = tw[((size_t)(t * nsm + ks) * 9 + l) * nbp + m];
    acc = fma_s(w * cs, u[((size_t)r * 9 + op.pi[g * 9 + l]) * nbp + m], acc);
  }
  cpr[(size_t)tl * nbp + m] = sc ? bilin(acc, inv(sc[m])) : acc;  // the vector is s_k u_k under lagged normalisation
}

// Full reorthogonalisation of the real chains (lanczos with reorth, src/reference/asr_recursion.f90:44-47):
// w = w - Q (Q^T w) against the stored vectors q_0..q_{nq-1} of every member (Q: nq vectors of qstride values).
// k_qdots: per block partials of <q_j, w> for every j (plane j of part); k_qfin: their fixed-order sums (one block per
// member); k_qsub: w -= sum_j c_j q_j, optionally with the partial sums w w of the result.
__global__ void __launch_bounds__(BLOCK) k_qdots(OpDev op, MemDev md, const double* w, const double* Q, size_t qstride, int nq,
                                                 double* part, size_t plane) {
  __shared__ double red[BLOCK];
  if (md.tact && !md.tact[blockIdx.y]) return;
  const int mt = md.mt, npw = WAVE / mt;
  const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x >> 6;
  const int m = blockIdx.y * mt + (lane & (mt - 1));
  const int i = (xcd_block(blockIdx.x, gridDim.x) * (BLOCK / WAVE) + wave) * npw + lane / mt;
  const bool live = i < op.nn && m < md.nb;
  double x[9];
  for (int l = 0; l < 9; ++l) x[l] = live ? w[((size_t)i * 9 + l) * md.nbp + m] : 0.0;
  for (int j = 0; j < nq; ++j) {
    double d = 0.0;
    if (live) {
      const double* q = Q + (size_t)j * qstride;
      for (int l = 0; l < 9; ++l) d = fma(q[((size_t)i * 9 + l) * md.nbp + m], x[l], d);
    }
    block_reduce_member<double>(d, mt, red, part + (size_t)j * plane + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * mt +
                                                threadIdx.x);
    __syncthreads();
  }
}

__global__ void __launch_bounds__(BLOCK) k_qfin(const double* part, size_t plane, int nbx, int mt, int nq, int nbp,
                                                double* coef) {
  __shared__ double lds[BLOCK / WAVE];
  const int m = blockIdx.x;
  for (int j = 0; j < nq; ++j) {
    const double v = block_sum_partials<double>(part + (size_t)j * plane, nbx, mt, m, lds);
    if (threadIdx.x == 0) coef[(size_t)j * nbp + m] = v;
    __syncthreads();
  }
}

__global__ void __launch_bounds__(BLOCK) k_qsub(OpDev op, MemDev md, double* w, const double* Q, size_t qstride, int nq,
                                                const double* coef, double* partial) {
  __shared__ double red[BLOCK];
  if (md.tact && !md.tact[blockIdx.y]) return;
  const int mt = md.mt, npw = WAVE / mt;
  const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x >> 6;
  const int m = blockIdx.y * mt + (lane & (mt - 1));
  const int i = (xcd_block(blockIdx.x, gridDim.x) * (BLOCK / WAVE) + wave) * npw + lane / mt;
  double dv = 0.0;
  if (i < op.nn && m < md.nb) {
    double t[9];
    for (int l = 0; l < 9; ++l) t[l] = 0.0;
    for (int j = 0; j < nq; ++j) {
      const double c = coef[(size_t)j * md.nbp + m];
      const double* q = Q + (size_t)j * qstride;
      for (int l = 0; l < 9; ++l) t[l] = fma(q[((size_t)i * 9 + l) * md.nbp + m], c, t[l]);
    }
    for (int l = 0; l < 9; ++l) {
      const size_t q = ((size_t)i * 9 + l) * md.nbp + m;
      const double v = w[q] - t[l];
      w[q] = v;
      dv = fma(v, v, dv);
    }
  }
  if (partial) block_reduce_member<double>(dv, mt, red, partial + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * mt + threadIdx.x);
}

// u <- w * inv_b (in place), members fastest
template <class T>
__global__ void k_scale(T* w, const T* inv_b, size_t n, int nbp) {
  size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  int m = (int)(q % nbp);
  w[q] = bilin(inv_b[m], w[q]);
}

// ------------------------------------------------------------------------------------------------ KPM (supercell)
struct RealDev {
  int nsite, nslot, nsp;
  const int* nbr;               // nsite x nslot
  const unsigned char* spec;    // nsite
  const double* par;            // 4 x nsp x 9: c, d, o, e
  const double* blk;            // (nslot + 1) x 81, slot 0 = on-site
  const double* idv;            // nsp x 9: 1 / d (no division in the steps)
  const double* fsh;            // nsp x 9: (c - cshift) / d^2 of the current recurrence's shift
};

__device__ inline uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ inline double rademacher(uint64_t seed, uint64_t v, uint64_t row) {
  return (splitmix64((seed << 40) ^ (v << 28) ^ row) >> 63) ? -1.0 : 1.0;
}

// Value of the input vector: x stored, or the Rademacher vector itself (x == null).
__device__ inline double kpm_in(const double* x, int nvp, uint64_t seed, int site, int l, int v) {
  return x ? x[((size_t)site * 9 + l) * nvp + v] : rademacher(seed, (uint64_t)v, (uint64_t)site * 9 + l);
}

// One Chebyshev step for a tile of VT vectors: y = alpha (h x - c x) / s - beta t0  (h = c + d S d, order 2 with o = 0),
// partial mu: sum_l r_l y_l.  Lanes = (site, vector) with VT vectors per site; a wave covers 64 / VT sites, their
// neighbour indices and bond blocks are wave-uniform for VT = 64.  The 19 blocks sit in LDS.  x == null / t0 == null: the
// Rademacher vector (generated, never stored).
template <int VT>
__global__ void __launch_bounds__(BLOCK) k_kpm_step(RealDev R, int nvec, int nvp, const double* __restrict__ x,
                                                    const double* __restrict__ t0, double* __restrict__ y, double cshift,
                                                    double alpha_over_s, double beta, uint64_t seed, double* partial,
                                                    int want_dot) {
  __shared__ double blk[19 * 81 + 64];
  __shared__ double red[BLOCK];
  const int nb = (R.nslot + 1) * 81;
  for (int q = threadIdx.x; q < nb; q += BLOCK) blk[q] = R.blk[q];
  __syncthreads();
  constexpr int SPW = WAVE / VT;  // sites per wave
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int v = blockIdx.y * VT + (lane % VT);
  const int site = (lb * (BLOCK / WAVE) + wave) * SPW + lane / VT;
  double dot = 0.0;
  if (site < R.nsite && v < nvec) {
    const int sp = R.spec[site];
    const double* cpar = R.par + (size_t)sp * 9;
    const double* dpar = R.par + (size_t)(R.nsp + sp) * 9;
    double acc[9], xv[9];
    for (int l = 0; l < 9; ++l) {
      xv[l] = kpm_in(x, nvp, seed, site, l, v);
      acc[l] = 0.0;
    }
    {
      double t[9];
      for (int l = 0; l < 9; ++l) t[l] = dpar[l] * xv[l];
      for (int l = 0; l < 9; ++l) {
        double a = acc[l];
        for (int l2 = 0; l2 < 9; ++l2) a = fma(blk[l * 9 + l2], t[l2], a);
        acc[l] = a;
      }
    }
    const int* nb_ = R.nbr + (size_t)site * R.nslot;
    for (int k = 0; k < R.nslot; ++k) {
      const int j = nb_[k];
      const double* dj = R.par + (size_t)(R.nsp + R.spec[j]) * 9;
      double t[9];
      for (int l = 0; l < 9; ++l) t[l] = dj[l] * kpm_in(x, nvp, seed, j, l, v);
      const double* bk = blk + (k + 1) * 81;
      for (int l = 0; l < 9; ++l) {
        double a = acc[l];
        for (int l2 = 0; l2 < 9; ++l2) a = fma(bk[l * 9 + l2], t[l2], a);
        acc[l] = a;
      }
    }
    for (int l = 0; l < 9; ++l) {
      double hx = cpar[l] * xv[l] + dpar[l] * acc[l];
      double yy = alpha_over_s * (hx - cshift * xv[l]);
      if (beta != 0.0) yy -= beta * kpm_in(t0, nvp, seed, site, l, v);
      y[((size_t)site * 9 + l) * nvp + v] = yy;
      if (want_dot) dot = fma(rademacher(seed, (uint64_t)v, (uint64_t)site * 9 + l), yy, dot);
    }
  }
  if (want_dot) {
    // per (block, vector): the sites of the wave by a shuffle tree, then the 4 waves in order
    for (int off = VT; off < WAVE; off <<= 1) dot += __shfl_xor(dot, off);
    if (lane < VT) red[wave * VT + lane] = dot;
    __syncthreads();
    if (threadIdx.x < VT) {
      double s = red[threadIdx.x];
      for (int w = 1; w < BLOCK / WAVE; ++w) s += red[w * VT + threadIdx.x];
      partial[((size_t)blockIdx.y * gridDim.x + lb) * VT + threadIdx.x] = s;
    }
  }
}

// Chebyshev step on d-scaled vectors, one site per wave (wave-uniform site: its neighbour list, species and the bond blocks
// are read through the scalar cache), lanes = 64 vectors.  Stored vectors are u_n = d o t_n, so the gathered neighbour
// values enter the bond blocks directly (h t = c t + d S (d t) needs d o t at the neighbours):
//     t2 = alpha (c t1 + d (B0 u1_i + sum_k B_k u1_j) - cshift t1) - beta t0,   t1 = u1 / d,  u2 = d o t2,
// partial mu: sum_l r_l t2_l.  GEN: t1 = r (the Rademacher vector, generated) for the first steps.  The neighbour loads
// of slot k + 1 are issued before the block of slot k is applied.
template <bool GEN, int NSLOT>
__global__ void __launch_bounds__(BLOCK) k_kpm_u(RealDev R, const double* __restrict__ blk, int nvec, int nvp,
                                                 const double* __restrict__ u1, const double* __restrict__ u0,
                                                 double* __restrict__ u2, double cshift, double alpha, double beta,
                                                 uint64_t seed, double* partial, int want_dot) {
  __shared__ double red[BLOCK];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int site = __builtin_amdgcn_readfirstlane(lb * (BLOCK / WAVE) + wave);
  const int v = blockIdx.y * 64 + lane;
  double dot = 0.0;
  if (site < R.nsite) {
    const int sp = R.spec[site];
    const double* cp = R.par + (size_t)sp * 9;
    const double* dp = R.par + (size_t)(R.nsp + sp) * 9;
    const double* id = R.idv + (size_t)sp * 9;
    const int nslot = NSLOT > 0 ? NSLOT : R.nslot;
    double acc[9], xo[9];
    {
      const double* __restrict__ row = u1 + (size_t)site * 9 * nvp;
      for (int l = 0; l < 9; ++l)
        xo[l] = GEN ? dp[l] * rademacher(seed, (uint64_t)v, (uint64_t)site * 9 + l) : row[(size_t)l * nvp + v];
    }
    for (int l = 0; l < 9; ++l) {
      double a = 0.0;
      for (int l2 = 0; l2 < 9; ++l2) a = fma(blk[l * 9 + l2], xo[l2], a);
      acc[l] = a;
    }
    const int* __restrict__ nbl = R.nbr + (size_t)site * R.nslot;
#pragma unroll
    for (int k = 0; k < (NSLOT > 0 ? NSLOT : 1); ++k) {
      for (int kk = (NSLOT > 0 ? k : 0); kk < (NSLOT > 0 ? k + 1 : nslot); ++kk) {
        const int j = nbl[kk];
        double xn[9];
        if (GEN) {
          const double* dj = R.par + (size_t)(R.nsp + R.spec[j]) * 9;
          for (int l = 0; l < 9; ++l) xn[l] = dj[l] * rademacher(seed, (uint64_t)v, (uint64_t)j * 9 + l);
        } else {
          const double* __restrict__ row = u1 + (size_t)j * 9 * nvp;
          for (int l = 0; l < 9; ++l) xn[l] = row[(size_t)l * nvp + v];
        }
        const double* __restrict__ bk = blk + (size_t)(kk + 1) * 81;
        for (int l = 0; l < 9; ++l) {
          double a = acc[l];
          for (int l2 = 0; l2 < 9; ++l2) a = fma(bk[l * 9 + l2], xn[l2], a);
          acc[l] = a;
        }
      }
    }
    if (v < nvec) {
      for (int l = 0; l < 9; ++l) {
        const uint64_t row = (uint64_t)site * 9 + l;
        const double r = rademacher(seed, (uint64_t)v, row);
        const double t1 = GEN ? r : xo[l] * id[l];
        const double hx = cp[l] * t1 + dp[l] * acc[l];
        double t2 = alpha * (hx - cshift * t1);
        if (beta != 0.0) t2 -= beta * (u0 ? u0[row * nvp + v] * id[l] : r);
        u2[row * nvp + v] = dp[l] * t2;
        if (want_dot) dot = fma(r, t2, dot);
      }
    }
  }
  if (want_dot) {
    red[threadIdx.x] = dot;
    __syncthreads();
    if (threadIdx.x < 64) {
      double s2 = red[threadIdx.x];
      for (int w = 1; w < BLOCK / WAVE; ++w) s2 += red[w * 64 + threadIdx.x];
      partial[((size_t)blockIdx.y * gridDim.x + lb) * 64 + threadIdx.x] = s2;
    }
  }
}

// The same step with two vectors per lane (16-byte loads: half the L1 requests of one vector per lane) and two sites per
// wave (32 lanes x 2 vectors = the 64-vector tile of one site).  DBL (moment doubling): instead of <r|t_{n+1}> the partials
// of <t_{n+1}|t_{n+1}> (partial) and <t_{n+1}|t_n> (partial2), from which mu_{2n+2} and mu_{2n+1} follow; no random
// vector is regenerated.
struct d2 {
  double x, y;
};
// fixed-order sum of a (vector pair) value over the two sites of each wave and the 4 waves; 64 values out per block
__device__ inline void kpm_pair_reduce(d2 v, d2 (*red)[32], int wave, int half, int vl, double* out) {
  v.x += __shfl_xor(v.x, 32);
  v.y += __shfl_xor(v.y, 32);
  if (half == 0) red[wave][vl] = v;
  __syncthreads();
  if (threadIdx.x < 32) {
    d2 s2 = red[0][threadIdx.x];
    for (int w = 1; w < BLOCK / WAVE; ++w) { s2.x += red[w][threadIdx.x].x; s2.y += red[w][threadIdx.x].y; }
    out[2 * threadIdx.x] = s2.x;
    out[2 * threadIdx.x + 1] = s2.y;
  }
}

// partials of <t|t> of a stored (d-scaled) vector, laid out as k_kpm_v2's (moment doubling: mu_2 = 2 <t_1|t_1> - mu_0)
__global__ void __launch_bounds__(BLOCK) k_kpm_sq(RealDev R, int nvp, const double* __restrict__ u, double* partial) {
  __shared__ d2 red[BLOCK / WAVE][32];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, half = lane >> 5, vl = lane & 31;
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int site = (lb * (BLOCK / WAVE) + wave) * 2 + half;
  const int v = blockIdx.y * 64 + 2 * vl;
  d2 dot = {0.0, 0.0};
  if (site < R.nsite) {
    const double* id = R.idv + (size_t)R.spec[site] * 9;
    for (int l = 0; l < 9; ++l) {
      const d2 x = ((const d2*)(u + ((size_t)site * 9 + l) * nvp + v))[0];
      const double tx = x.x * id[l], ty = x.y * id[l];
      dot.x = fma(tx, tx, dot.x);
      dot.y = fma(ty, ty, dot.y);
    }
  }
  kpm_pair_reduce(dot, red, wave, half, vl, partial + ((size_t)blockIdx.y * gridDim.x + lb) * 64);
}

template <int NSLOT, bool DBL>
__global__ void __launch_bounds__(BLOCK) k_kpm_v2(RealDev R, const double* __restrict__ blk, int nvec, int nvp,
                                                  const double* __restrict__ u1, const double* __restrict__ u0,
                                                  double* __restrict__ u2, double cshift, double alpha, double beta,
                                                  uint64_t seed, double* partial, double* partial2) {
  __shared__ d2 red[BLOCK / WAVE][32];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, half = lane >> 5, vl = lane & 31;
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int site = (lb * (BLOCK / WAVE) + wave) * 2 + half;
  const int v = blockIdx.y * 64 + 2 * vl;
  d2 dot = {0.0, 0.0}, dot2 = {0.0, 0.0};
  if (site < R.nsite) {
    const int sp = R.spec[site];
    const double* dp = R.par + (size_t)(R.nsp + sp) * 9;
    const double* id = R.idv + (size_t)sp * 9;
    const double* fs = R.fsh + (size_t)sp * 9;
    const int nslot = NSLOT > 0 ? NSLOT : R.nslot;
    // acc = block0 x + ((c - cshift) / d^2) x: the own-row term of t2 folded into the start of the sum (x = d t1 is
    // the stored row), so that the row is dead after this and does not hold 18 registers through the neighbour loop
    d2 acc[9];
    {
      const d2* __restrict__ row = (const d2*)(u1 + (size_t)site * 9 * nvp + v);
      d2 xo[9];
      for (int l = 0; l < 9; ++l) xo[l] = row[(size_t)l * (nvp / 2)];
      for (int l = 0; l < 9; ++l) {
        const double f = fs[l];
        d2 a = {f * xo[l].x, f * xo[l].y};
        for (int l2 = 0; l2 < 9; ++l2) {
          a.x = fma(blk[l * 9 + l2], xo[l2].x, a.x);
          a.y = fma(blk[l * 9 + l2], xo[l2].y, a.y);
        }
        acc[l] = a;
      }
    }
    const int* __restrict__ nbl = R.nbr + (size_t)site * R.nslot;
#pragma unroll 2
    for (int kk = 0; kk < nslot; ++kk) {
      const int j = nbl[kk];
      const d2* __restrict__ row = (const d2*)(u1 + (size_t)j * 9 * nvp + v);
      d2 xn[9];
      for (int l = 0; l < 9; ++l) xn[l] = row[(size_t)l * (nvp / 2)];
      const double* __restrict__ bk = blk + (size_t)(kk + 1) * 81;
      for (int l = 0; l < 9; ++l) {
        d2 a = acc[l];
        for (int l2 = 0; l2 < 9; ++l2) {
          a.x = fma(bk[l * 9 + l2], xn[l2].x, a.x);
          a.y = fma(bk[l * 9 + l2], xn[l2].y, a.y);
        }
        acc[l] = a;
      }
    }
    // t_n-1 for the nine orbitals before the first store (a load behind a store would wait for it as well)
    d2 t0a[9];
#pragma unroll
    for (int l = 0; l < 9; ++l) {
      const nt_d2 t0n = __builtin_nontemporal_load((const nt_d2*)(u0 + ((uint64_t)site * 9 + l) * nvp + v));
      t0a[l] = d2{t0n.x, t0n.y};
    }
    d2 xoa[DBL ? 9 : 1];
    if constexpr (DBL) {
#pragma unroll
      for (int l = 0; l < 9; ++l) xoa[l] = ((const d2*)(u1 + ((uint64_t)site * 9 + l) * nvp + v))[0];
    }
    for (int l = 0; l < 9; ++l) {
      const uint64_t row = (uint64_t)site * 9 + l;
      const d2 t0v = t0a[l];
      double t2x = alpha * (dp[l] * acc[l].x) - beta * (t0v.x * id[l]);
      double t2y = alpha * (dp[l] * acc[l].y) - beta * (t0v.y * id[l]);
      if (v + 1 < nvec) {
        __builtin_nontemporal_store(nt_d2{dp[l] * t2x, dp[l] * t2y}, (nt_d2*)(u2 + row * nvp + v));
      } else if (v < nvec) {
        u2[row * nvp + v] = dp[l] * t2x;
      }
      if constexpr (DBL) {
        // the own row again (an L1/L2 hit) rather than held through the neighbour loop
        const d2 xo = xoa[l];
        dot.x = fma(t2x, t2x, dot.x);
        dot.y = fma(t2y, t2y, dot.y);
        dot2.x = fma(t2x, xo.x * id[l], dot2.x);
        dot2.y = fma(t2y, xo.y * id[l], dot2.y);
      } else {
        const double r0 = rademacher(seed, (uint64_t)v, row), r1 = rademacher(seed, (uint64_t)v + 1, row);
        dot.x = fma(r0, t2x, dot.x);
        dot.y = fma(r1, t2y, dot.y);
      }
    }
  }
  // per (block, vector): the two sites of the wave, then the 4 waves in order
  kpm_pair_reduce(dot, red, wave, half, vl, partial + ((size_t)blockIdx.y * gridDim.x + lb) * 64);
  if constexpr (DBL) {
    __syncthreads();
    kpm_pair_reduce(dot2, red, wave, half, vl, partial2 + ((size_t)blockIdx.y * gridDim.x + lb) * 64);
  }
}

// The Chebyshev step with the block contraction on the matrix cores (FP64 MFMA 16x16x4): one wave per site and 64 vectors
// (four 16-vector tiles).  The contraction Y = sum_k B_k X_nbr(k) over the on-site block and the NSLOT bond blocks is one
// product of a 16 x 171 matrix (the nine output orbitals, padded; the 19 blocks side by side) by a 171 x 16 matrix per
// tile (the gathered rows stacked), in 43 steps of k = 4: the A operand (blocks, the same for every site) from LDS, the
// B operand one gathered value per lane (rows of the neighbour in slot (4 s + g) / 9, orbital (4 s + g) % 9 for the lane
// group g = lane / 16), so a k-step mixes the end of one block with the start of the next and no k is padded but the last
// one.  The MFMA result holds D[g + 4 r][lane % 16] in register r.  The own-row term ((c - cshift) / d^2) x, the
// recurrence, the stores and the moment partials follow per lane for its rows (g, g + 4, and 8 in group 0), as k_kpm_v2.
typedef double kd4 __attribute__((ext_vector_type(4)));
template <int NSLOT, bool DBL>
__global__ void __launch_bounds__(BLOCK) k_kpm_mf(RealDev R, const double* __restrict__ blk, int nvec, int nvp,
                                                  const double* __restrict__ u1, const double* __restrict__ u0,
                                                  double* __restrict__ u2, double cshift, double alpha, double beta,
                                                  uint64_t seed, double* partial, double* partial2) {
  constexpr int KT = (NSLOT + 1) * 9;  // 171 for 18 slots
  constexpr int NKS = (KT + 3) / 4;    // 43 k-steps
  __shared__ double at[NKS * WAVE];    // A operand of k-step s for lane l at s * 64 + l
  __shared__ int roff[BLOCK / WAVE][NSLOT + 1];
  __shared__ double red[BLOCK / WAVE][WAVE];
  for (int q = threadIdx.x; q < NKS * WAVE; q += BLOCK) {
    const int st = q / WAVE, l = q % WAVE, row = l % 16, kk = 4 * st + l / 16;
    at[q] = (row < 9 && kk < KT) ? blk[(kk / 9) * 81 + row * 9 + kk % 9] : 0.0;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4, col = lane & 15;
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int site = lb * (BLOCK / WAVE) + wave;
  const bool live = site < R.nsite;
  if (lane <= NSLOT) roff[wave][lane] = live ? (lane == 0 ? site : R.nbr[(size_t)site * R.nslot + lane - 1]) * 9 : 0;
  __syncthreads();
  const int v0 = blockIdx.y * 64;
  double dsum[4] = {0.0, 0.0, 0.0, 0.0}, dsum2[4] = {0.0, 0.0, 0.0, 0.0};
  if (live) {
    kd4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = kd4{0.0, 0.0, 0.0, 0.0};
    const double* __restrict__ xb = u1 + v0 + col;
#pragma unroll 4
    for (int st = 0; st < NKS; ++st) {
      const int kk = 4 * st + g;
      const bool ok = kk < KT;
      const size_t r = (size_t)(roff[wave][ok ? kk / 9 : 0] + (ok ? kk % 9 : 0)) * nvp;
      const double a = at[st * WAVE + lane];
      double b[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) b[t] = ok ? xb[r + 16 * t] : 0.0;
#pragma unroll
      for (int t = 0; t < 4; ++t) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b[t], acc[t], 0, 0, 0);
    }
    const int sp = R.spec[site];
    const double* dp = R.par + (size_t)(R.nsp + sp) * 9;
    const double* id = R.idv + (size_t)sp * 9;
    const double* fs = R.fsh + (size_t)sp * 9;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int v = v0 + 16 * t + col;
#pragma unroll
      for (int rr = 0; rr < 3; ++rr) {
        const int l = g + 4 * rr;
        if (l >= 9 || v >= nvec) continue;
        const uint64_t row = (uint64_t)site * 9 + l;
        const double xo = u1[row * nvp + v];
        const double hsum = acc[t][rr] + fs[l] * xo;
        const double t2 = alpha * (dp[l] * hsum) - beta * (__builtin_nontemporal_load(u0 + row * nvp + v) * id[l]);
        __builtin_nontemporal_store(dp[l] * t2, u2 + row * nvp + v);
        if constexpr (DBL) {
          dsum[t] = fma(t2, t2, dsum[t]);
          dsum2[t] = fma(t2, xo * id[l], dsum2[t]);
        } else {
          dsum[t] = fma(rademacher(seed, (uint64_t)v, row), t2, dsum[t]);
        }
      }
    }
  }
  // per vector: the four lane groups, then the four waves (sites) of the block in order
  auto reduce = [&](double (&ds)[4], double* out) {
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      double vsum = ds[t];
      vsum += __shfl_xor(vsum, 16);
      vsum += __shfl_xor(vsum, 32);
      if (g == 0) red[wave][16 * t + col] = vsum;
    }
    __syncthreads();
    if (threadIdx.x < WAVE) {
      double s2 = red[0][threadIdx.x];
      for (int w = 1; w < BLOCK / WAVE; ++w) s2 += red[w][threadIdx.x];
      out[((size_t)blockIdx.y * gridDim.x + lb) * 64 + threadIdx.x] = s2;
    }
  };
  reduce(dsum, partial);
  if constexpr (DBL) {
    __syncthreads();
    reduce(dsum2, partial2);
  }
}

// The Chebyshev step on vector slices: the 64-vector rows are processed as slices of 16 vectors (one 128-byte line per
// site and orbital), 8 lanes per site with two vectors each, 8 sites per wave; blockIdx.y is the slice, so one slice of
// all sites is swept before the next and the slice of the input (a quarter of the vector) stays in the last-level caches
// while its rows are gathered 19 times.  Bond blocks from LDS, one row at a time.  Same arithmetic and partials layout as
// k_kpm_v2 (partials indexed by vector within its 64-vector tile).
template <int NSLOT>
__global__ void __launch_bounds__(BLOCK) k_kpm_s16(RealDev R, int nvec, int nvp, const double* __restrict__ u1,
                                                   const double* __restrict__ u0, double* __restrict__ u2, double cshift,
                                                   double alpha, double beta, uint64_t seed, double* partial) {
  extern __shared__ double bls[];
  __shared__ d2 red[BLOCK / WAVE][8][8];
  const int nb = (NSLOT + 1) * 81;
  for (int q = threadIdx.x; q < (NSLOT + 1) * 90; q += BLOCK) bls[q] = (q % 10 < 9) ? R.blk[(q / 90) * 81 + ((q % 90) / 10) * 9 + q % 10] : 0.0;
  (void)nb;
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, sl = lane >> 3, vp = lane & 7;
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int site = (lb * (BLOCK / WAVE) + wave) * 8 + sl;
  const int v = blockIdx.y * 16 + 2 * vp;
  d2 dot = {0.0, 0.0};
  if (site < R.nsite) {
    const int sp = R.spec[site];
    const double* cp = R.par + (size_t)sp * 9;
    const double* dp = R.par + (size_t)(R.nsp + sp) * 9;
    const double* id = R.idv + (size_t)sp * 9;
    d2 acc[9], xo[9];
    for (int l = 0; l < 9; ++l) acc[l] = d2{0.0, 0.0};
    const int* __restrict__ nbl = R.nbr + (size_t)site * NSLOT;
#pragma unroll 1
    for (int k = 0; k <= NSLOT; ++k) {
      const int j = k == 0 ? site : nbl[k - 1];
      const d2* __restrict__ row = (const d2*)(u1 + (size_t)j * 9 * nvp + v);
      d2 xn[9];
#pragma unroll
      for (int l = 0; l < 9; ++l) xn[l] = row[(size_t)l * (nvp / 2)];
      if (k == 0)
        for (int l = 0; l < 9; ++l) xo[l] = xn[l];
      const double* __restrict__ bk = bls + k * 90;
#pragma unroll
      for (int l = 0; l < 9; ++l) {
        d2 a = acc[l];
#pragma unroll
        for (int l2 = 0; l2 < 9; ++l2) {
          a.x = fma(bk[l * 10 + l2], xn[l2].x, a.x);
          a.y = fma(bk[l * 10 + l2], xn[l2].y, a.y);
        }
        acc[l] = a;
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    for (int l = 0; l < 9; ++l) {
      const uint64_t row = (uint64_t)site * 9 + l;
      const double r0 = rademacher(seed, (uint64_t)v, row), r1 = rademacher(seed, (uint64_t)v + 1, row);
      const d2 t0v = ((const d2*)(u0 + row * nvp + v))[0];
      const double t1x = xo[l].x * id[l], t1y = xo[l].y * id[l];
      double t2x = alpha * (cp[l] * t1x + dp[l] * acc[l].x - cshift * t1x) - beta * (t0v.x * id[l]);
      double t2y = alpha * (cp[l] * t1y + dp[l] * acc[l].y - cshift * t1y) - beta * (t0v.y * id[l]);
      if (v < nvec) u2[row * nvp + v] = dp[l] * t2x;
      if (v + 1 < nvec) u2[row * nvp + v + 1] = dp[l] * t2y;
      dot.x = fma(r0, t2x, dot.x);
      dot.y = fma(r1, t2y, dot.y);
    }
  }
  // per (block, vector): the 8 sites of the wave, then the 4 waves in order
  red[wave][sl][vp] = dot;
  __syncthreads();
  if (threadIdx.x < 8) {
    d2 s2 = {0.0, 0.0};
    for (int w = 0; w < BLOCK / WAVE; ++w)
      for (int q = 0; q < 8; ++q) { s2.x += red[w][q][threadIdx.x].x; s2.y += red[w][q][threadIdx.x].y; }
    const int vv = blockIdx.y * 16 + 2 * threadIdx.x;
    double* pp = partial + ((size_t)(vv / 64) * gridDim.x + lb) * 64 + (vv % 64);
    pp[0] = s2.x;
    pp[1] = s2.y;
  }
}

// mu_per[v][n] = (1/dim) sum_blocks partial, one block per vector, fixed order
// Lanczos on the supercell (spectral_bounds): nvp columns of nrow rows, column v = chain v.  Column sums of x .* y per
// block (rows strided over the grid, fixed order), summed on the host.
__global__ void __launch_bounds__(BLOCK) k_col_dot(const double* __restrict__ x, const double* __restrict__ y, size_t nrow,
                                                   int nvp, double* partial) {
  __shared__ double red[BLOCK];
  const int v = threadIdx.x % nvp, rpb = BLOCK / nvp;
  double s = 0.0;
  for (size_t r = (size_t)blockIdx.x * rpb + threadIdx.x / nvp; r < nrow; r += (size_t)gridDim.x * rpb)
    s = fma(x[r * nvp + v], y[r * nvp + v], s);
  red[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < (unsigned)nvp) {
    double t = 0.0;
    for (int q = threadIdx.x; q < BLOCK; q += nvp) t += red[q];
    partial[(size_t)blockIdx.x * nvp + threadIdx.x] = t;
  }
}

// random start columns: u(row, v) = U[-1/2, 1/2) from splitmix64(((seed + v) << 40) ^ row) (columns v >= nvec zero)
__global__ void k_rl_start(double* u, size_t nrow, int nvp, int nvec, uint64_t seed) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nrow * nvp) return;
  const int v = (int)(q % nvp);
  const size_t r = q / nvp;
  u[q] = v < nvec ? (double)(splitmix64(((seed + (uint64_t)v) << 40) ^ r) >> 11) * 0x1.0p-53 - 0.5 : 0.0;
}

// w = w - a u - b up per column (coef: a then b, nvp each)
__global__ void k_rl_axpy(double* w, const double* u, const double* up, const double* coef, size_t n, int nvp) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int v = (int)(q % nvp);
  w[q] = w[q] - coef[v] * u[q] - coef[nvp + v] * up[q];
}

// up = u; u = w * s per column (s: nvp); w == nullptr: u = u * s
__global__ void k_rl_shift(double* u, double* up, const double* w, const double* sc, size_t n, int nvp) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int v = (int)(q % nvp);
  if (w) {
    up[q] = u[q];
    u[q] = w[q] * sc[v];
  } else {
    u[q] = u[q] * sc[v];
  }
}

// The moments' partial sums in two stages (instead of one block per vector, whose strided reads fetched each line once
// per vector: 73 us per step, 146 with the doubling's two sums, against the 1.65 ms step): k_kpm_psum, block (g, tv) of
// KPS x ntv with KPW waves, wave w summing the rows b = g + KPS (KPW r + w), r = 0, 1, ... of vector tile tv in order
// (lanes = the tile's vectors, each row read whole), then the waves in order; k_kpm_pfin, one block per vector tile,
// sums the KPS values of each vector, KPS / KPF_WAVES in order per wave, then the waves in order.  part2 (moment
// doubling): the second partials alike, after the first in stage.
constexpr int KPS = 256, KPW = 4, KPF_WAVES = 16;
__global__ void __launch_bounds__(WAVE * KPW) k_kpm_psum(const double* __restrict__ part, const double* __restrict__ part2,
                                                         int nbx, int vt, double* __restrict__ stage) {
  __shared__ double red[2][KPW][WAVE];
  const int g = blockIdx.x, tv = blockIdx.y, lane = threadIdx.x & (WAVE - 1), w = threadIdx.x >> 6;
  double s = 0.0, s2 = 0.0;
  if (lane < vt) {
    // the loads of several rows in flight ahead of the (ordered) sum
    const int b0 = g + KPS * w, bs = KPS * KPW;
    const int nr = b0 < nbx ? (nbx - 1 - b0) / bs + 1 : 0;
    const size_t o0 = ((size_t)tv * nbx + b0) * vt + lane, st = (size_t)bs * vt;
#pragma unroll 8
    for (int r = 0; r < nr; ++r) s += part[o0 + r * st];
    if (part2) {
#pragma unroll 8
      for (int r = 0; r < nr; ++r) s2 += part2[o0 + r * st];
    }
  }
  red[0][w][lane] = s;
  red[1][w][lane] = s2;
  __syncthreads();
  if (w != 0 || lane >= vt) return;
  s = red[0][0][lane];
  s2 = red[1][0][lane];
  for (int k = 1; k < KPW; ++k) {
    s += red[0][k][lane];
    s2 += red[1][k][lane];
  }
  const size_t so = ((size_t)tv * KPS + g) * vt + lane;
  stage[so] = s;
  if (part2) stage[(size_t)gridDim.y * KPS * vt + so] = s2;
}
// m < 0: mu_per[v][n] = sum / dim; m >= 1 (moment doubling, T_2m = 2 T_m^2 - T_0, T_2m-1 = 2 T_m T_m-1 - T_1):
// mu_2m = 2 <t_m|t_m> / dim - mu_0 and, with the second partials (two), mu_2m-1 = 2 <t_m|t_m-1> / dim - mu_1
__global__ void __launch_bounds__(WAVE * KPF_WAVES) k_kpm_pfin(const double* __restrict__ stage, int two, int vt, int ntv,
                                                               int nvec, double inv_dim, double* mu_per, int nmom, int n,
                                                               int m) {
  __shared__ double red[2][KPF_WAVES][WAVE];
  const int tv = blockIdx.x, lane = threadIdx.x & (WAVE - 1), w = threadIdx.x >> 6;
  double s = 0.0, s2 = 0.0;
  if (lane < vt) {
    constexpr int per = KPS / KPF_WAVES;
    const double* __restrict__ st = stage + ((size_t)tv * KPS + w * per) * vt + lane;
#pragma unroll
    for (int g = 0; g < per; ++g) s += st[(size_t)g * vt];
    if (two) {
      const double* __restrict__ st2 = st + (size_t)ntv * KPS * vt;
#pragma unroll
      for (int g = 0; g < per; ++g) s2 += st2[(size_t)g * vt];
    }
  }
  red[0][w][lane] = s;
  red[1][w][lane] = s2;
  __syncthreads();
  const int v = tv * vt + lane;
  if (w != 0 || lane >= vt || v >= nvec) return;
  s = red[0][0][lane];
  s2 = red[1][0][lane];
  for (int k = 1; k < KPF_WAVES; ++k) {
    s += red[0][k][lane];
    s2 += red[1][k][lane];
  }
  double* mv = mu_per + (size_t)v * nmom;
  if (m < 0) {
    mv[n] = s * inv_dim;
  } else {
    if (2 * m < nmom) mv[2 * m] = 2.0 * s * inv_dim - mv[0];
    if (two && 2 * m - 1 < nmom) mv[2 * m - 1] = 2.0 * s2 * inv_dim - mv[1];
  }
}

// mu[n] = sum_v mu_per[v][n] / nvec, the vectors in order (the host's sum of asr_gpu_kpm_moments, on the device)
__global__ void k_kpm_mean(const double* __restrict__ mu_per, int nvec, int nmom, double* __restrict__ mu) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nmom) return;
  double acc = 0.0;
  for (int v = 0; v < nvec; ++v) acc += mu_per[(size_t)v * nmom + n];
  mu[n] = acc / (double)nvec;
}

}  // namespace

// ================================================================================================ host side
// Cached device blocks of the chain vectors: a batched run returns its level vectors here instead of freeing them, and the
// next run of the same operator (the next batch of a contour pass) takes them back without hipMalloc / hipFree, whose
// cost grows with the size (tens of GB per batch).  A free block serves a request of up to its size and at least 4/5 of
// it.  One host thread per operator (the calls of an operator are not concurrent), so no lock.
struct BlockPool {
  std::multimap<size_t, void*> free_;
  std::map<void*, size_t> size_;
  void* get(size_t bytes) {
    auto it = free_.lower_bound(bytes);
    if (it != free_.end() && it->first - bytes <= it->first / 5) {
      void* p = it->second;
      free_.erase(it);
      return p;
    }
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      trim();  // out of memory: give the cached blocks back and try once more
      if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    }
    size_[p] = bytes;
    return p;
  }
  void put(void* p) {
    auto it = size_.find(p);
    if (it == size_.end()) {
      (void)hipFree(p);
      return;
    }
    free_.insert({it->second, p});
  }
  void trim() {
    for (auto& kv : free_) {
      size_.erase(kv.second);
      (void)hipFree(kv.second);
    }
    free_.clear();
  }
  ~BlockPool() {
    for (auto& kv : size_) (void)hipFree(kv.first);
  }
};

struct asr_gpu_op {
  int nn, nslot, nop, nsmax, ncls, nsp, ntype, nstab;
  BlockPool pool;  // cached level-vector blocks of the batched runs (asr_gpu_op_trim releases them)
  std::vector<int> pi_h, nsc_h, spc_h;
  std::vector<double> sg_h, mt_h;  // mt: ncls x nsmax x nsmax x nsmax
  uint32_t *hop = nullptr, *flip = nullptr, *info = nullptr;
  double *sq = nullptr, *isq = nullptr, *sg = nullptr, *blk = nullptr, *blkp = nullptr;
  uint32_t* sgm = nullptr;
  int *stab_ptr = nullptr, *stab = nullptr, *pi = nullptr, *nsc = nullptr;
  int nplist = 0;                   // nodes with flip partners (INFO_PARTNERS) and their list on the device
  int* plist = nullptr;
  int ntarget = 0;                  // target states of the projected chains (asr_gpu_op_set_targets)
  int *tnode = nullptr, *tgop = nullptr;  // ntarget x nsmax: node (-1 none) and operation of target t in state k
  hipStream_t stream = nullptr;
  bool have_tables = false, have_ops = false, have_blocks = false, have_classes = false;
  bool piid = false;                              // every operation maps each orbital onto itself
  int tile_pref = 0;                              // member tile of the batched chains, 0 = autotuned
  double depth_tol = 0.0;                         // adaptive depth of the edc chains (ed_depth_tol), 0 = off
  int depth_tail = 0;                             // tail of that test (ed_tail): 0 hull, 1 bp, 2 avg2, 3 none
  bool reorth = false;                            // full reorthogonalisation of the real chains (lanczos reorth)
  bool compact = true;                            // member compaction between host-issued levels (adaptive depth)
  int window = 0;                                 // member slots of a window run (0: every chain holds a slot)
  bool queue_dev = true;                          // the window run's queue on the device (k_refill) while chains wait
  int ncompact = 0;                               // compactions of the last batched run
  int level_form = 0;                             // edc chains: 0 the fused level (a_k from the first pass, the update
                                                  // in the second), 1 the separate update pass; the adjoint's reverse
                                                  // sweep: 0 its combinations on the operator's last pass, 1 separate
  bool asym = false;                              // the fused level's check found the operator not symmetric: form 1
  bool slot_flippable[2] = {false, false};        // the site of frozen slot 1 / 2 was flippable (no species may be frozen)
  std::vector<std::pair<long long, int>> tuned;   // (members, complex, uniform d, zero o) key -> member tile
};

struct asr_gpu_real {
  int nsite, nslot, nsp, vt = 0;
  int device = 0;    // the device current at asr_gpu_real_create
  bool v2 = true;    // Chebyshev steps with two vectors per lane
  bool s16 = false;  // ... on 16-vector slices (8 sites per wave)
  bool dbl = false;  // kpm_moments by moment doubling (two moments per Chebyshev step)
  bool mf = false;   // ... the steps with the block contraction on the matrix cores (k_kpm_mf, 18 slots)
  int* nbr = nullptr;
  unsigned char* spec = nullptr;
  double *par = nullptr, *blk = nullptr;
  double* tab = nullptr;       // 2 x nsp x 9: 1 / d, then (c - cshift) / d^2 of the last recurrence
  std::vector<double> tab_h;   // its host copy
  std::vector<double> par_h;   // par on the host
  hipStream_t stream = nullptr;
};

namespace {

template <class P>
int dalloc(P** p, size_t n) {
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(P));
  if (e != hipSuccess) return fail(std::string("hipMalloc: ") + hipGetErrorString(e));
  return 0;
}

// a level-vector block from the operator's cache (returned with pool->put)
template <class P>
int palloc(BlockPool* pool, P** p, size_t n) {
  if (n == 0) n = 1;
  *p = (P*)pool->get(n * sizeof(P));
  if (!*p) return fail("hipMalloc: out of device memory for the level vectors");
  return 0;
}

OpDev dev_of(const asr_gpu_op* op) {
  OpDev d;
  d.nn = op->nn; d.nslot = op->nslot; d.nop = op->nop; d.nsmax = op->nsmax; d.ncls = op->ncls; d.nsp = op->nsp;
  d.ntype = op->ntype; d.nstab = op->nstab;
  d.hop = op->hop; d.flip = op->flip; d.info = op->info; d.sq = op->sq; d.stab_ptr = op->stab_ptr; d.stab = op->stab;
  d.isq = op->isq; d.sgm = op->sgm; d.blkp = op->blkp;
  d.pi = op->pi; d.sg = op->sg; d.blk = op->blk; d.nsc = op->nsc;
  d.piid = op->piid ? 1 : 0;
  return d;
}

int pow2_tile(int nb) {
  int t = 1;
  while (t < nb && t < 64) t <<= 1;
  return t;
}

// Per-member host preparation: characters, local weights, frozen rows, uniform d.
struct Members {
  int nb = 0, nbp = 0, mt = 1, pair = 0;
  bool uniform_d = true, zero_o = true, zero_e = true, complex_lin = false, ed3 = false;
  std::vector<uint64_t> chim;
  std::vector<double> wts, fpar, dvec, dperm, mtm;  // mtm: ncls x nsmax x nsmax x nsmax x nbp (class tables per member)
  std::vector<int> fsp;                             // 2 x nbp: frozen species of the start site and of the partner site
  std::vector<cplx> z;
};

// chi(g) of the start orbital from the monomial tables: D(g)[l0, l0] = sg(l0, g) if pi(l0, g) = l0
int chi_mask(const asr_gpu_op* op, int l0, uint64_t* mask) {
  uint64_t mk = 0;
  for (int g = 0; g < op->nop; ++g) {
    if (op->pi_h[g * 9 + l0] != l0) return fail("augcls_set_start: start orbital is not one-dimensional under the group");
    if (op->sg_h[g * 9 + l0] < 0) mk |= (1ull << g);
  }
  *mask = mk;
  return 0;
}

// params: nchain x 4 x nsp x 9 (c, d, o, e); pab: nchain x 2 x nsp x 9 or null; conc: nchain x ncls x nsmax or null;
// pex (route ed3): nchain x 4 x nsp x 9 (ua, ub, qa, qb) or null; frozen_sp2: species of the frozen partner site (nodes
// of fz slot 2, augcls_freeze_any) or null; chim_in: the characters of a symmetry-adapted start vector (bit g set <=>
// chi(g) = -1, augcls_set_chi) or null for those of the start orbital
int prepare_members(const asr_gpu_op* op, int nchain, const int* start_orb, const double* params, const double* pab,
                    const int* frozen_sp, const double* conc, const double* z, Members& M, int tile = 0,
                    const double* pex = nullptr, const int* frozen_sp2 = nullptr, const uint64_t* chim_in = nullptr) {
  const int nsm = op->nsmax, ncls = op->ncls, nsp = op->nsp;
  M.nb = nchain;
  M.pair = (tile == 65) ? 1 : 0;  // 65: member tile 64 with complex members as lane pairs
  if (tile == 65) tile = 64;
  M.mt = tile > 0 ? std::min(tile, pow2_tile(nchain)) : pow2_tile(nchain);
  if (M.mt != WAVE) M.pair = 0;
  M.nbp = ((nchain + M.mt - 1) / M.mt) * M.mt;
  const int nbp = M.nbp;
  M.chim.assign(nbp, 0);
  M.wts.assign((size_t)NQ * ncls * nsm * nsm * 9 * nbp, 0.0);
  M.fpar.assign((size_t)2 * NQ * 9 * nbp, 0.0);
  M.dvec.assign((size_t)9 * nbp, 0.0);
  M.dperm.assign((size_t)op->nop * 9 * nbp, 0.0);
  M.mtm.assign((size_t)ncls * nsm * nsm * nsm * nbp, 0.0);
  M.fsp.assign((size_t)2 * nbp, -1);
  M.z.assign(nbp, cplx{0.0, 0.0});
  M.complex_lin = (pab != nullptr);
  M.ed3 = (pex != nullptr);
  std::vector<double> mt((size_t)ncls * nsm * nsm * nsm);
  for (int m = 0; m < nchain; ++m) {
    if (start_orb[m] < 0 || start_orb[m] > 8) return fail("start orbital out of range");
    if (chim_in) {
      if (op->nop < 64 && (chim_in[m] >> op->nop) != 0) return fail("augcls_set_chi: character bits beyond the group");
      M.chim[m] = chim_in[m];
    } else if (chi_mask(op, start_orb[m], &M.chim[m])) {
      return 1;
    }
    // class tables: the operator's, or the member's concentrations (augcls_set_class_conc)
    if (conc) {
      std::fill(mt.begin(), mt.end(), 0.0);
      for (int kap = 0; kap < ncls; ++kap) {
        const double* cc = conc + ((size_t)m * ncls + kap) * nsm;
        int n = op->nsc_h[kap];
        double sum = 0.0;
        for (int s = 0; s < n; ++s) {
          if (cc[s] < 0.0) return fail("augcls_set_class_conc: concentrations");
          sum += cc[s];
        }
        if (std::fabs(sum - 1.0) > 1e-8) return fail("augcls_set_class_conc: concentrations");
        // augn_basis on the species of non-zero weight (class_mt, src/asr_augcls.f90:190-205)
        std::vector<int> sup;
        for (int s = 0; s < n; ++s)
          if (cc[s] > 0.0) sup.push_back(s);
        int ns = (int)sup.size();
        std::vector<double> phi((size_t)ns * ns, 0.0);
        for (int j = 0; j < ns; ++j) phi[j] = 1.0;
        for (int l = 1; l < ns; ++l) {
          for (int j = 0; j < ns; ++j) phi[l * ns + j] = 0.0;
          phi[l * ns + l - 1] = 1.0;
          for (int rep = 0; rep < 2; ++rep)
            for (int q = 0; q < l; ++q) {
              double ip = 0.0;
              for (int j = 0; j < ns; ++j) ip += cc[sup[j]] * phi[q * ns + j] * phi[l * ns + j];
              for (int j = 0; j < ns; ++j) phi[l * ns + j] -= ip * phi[q * ns + j];
            }
          double nr = 0.0;
          for (int j = 0; j < ns; ++j) nr += cc[sup[j]] * phi[l * ns + j] * phi[l * ns + j];
          nr = std::sqrt(nr);
          for (int j = 0; j < ns; ++j) phi[l * ns + j] /= nr;
        }
        for (int l = 0; l < ns; ++l)
          for (int k = 0; k < ns; ++k)
            for (int j = 0; j < ns; ++j)
              mt[(((size_t)kap * nsm + l) * nsm + k) * nsm + sup[j]] = cc[sup[j]] * phi[l * ns + j] * phi[k * ns + j];
      }
    } else {
      mt = op->mt_h;
    }
    const double* pm = params + (size_t)m * 4 * nsp * 9;
    for (int q = 0; q < NQ; ++q) {
      const double* par = nullptr;
      if (q < 4) par = pm + (size_t)q * nsp * 9;
      else if (q < 6 && pab) par = pab + ((size_t)m * 2 + (q - 4)) * nsp * 9;
      else if (q >= 6 && pex) par = pex + ((size_t)m * 4 + (q - 6)) * nsp * 9;
      if (!par) continue;
      // local weights wt(lam, k, kk, kap) = sum_s par(lam, spc(s, kap)) mt(k, kk, s, kap)  (local_weights)
      for (int kap = 0; kap < ncls; ++kap)
        for (int kk = 0; kk < op->nsc_h[kap]; ++kk)
          for (int k = 0; k < op->nsc_h[kap]; ++k)
            for (int lam = 0; lam < 9; ++lam) {
              double s = 0.0;
              for (int sp = 0; sp < op->nsc_h[kap]; ++sp)
                s += par[op->spc_h[kap * nsm + sp] * 9 + lam] * mt[(((size_t)kap * nsm + k) * nsm + kk) * nsm + sp];
              M.wts[(((((size_t)q * ncls + kap) * nsm + k) * nsm + kk) * 9 + lam) * nbp + m] = s;
            }
      for (int slot = 0; slot < 2; ++slot) {
        const int* fsp = slot == 0 ? frozen_sp : frozen_sp2;
        if (!fsp || fsp[m] < 0) continue;
        if (fsp[m] >= nsp) return fail("augcls_freeze_site: species index out of range");
        if (op->slot_flippable[slot]) return fail("augcls_freeze_site: the frozen site was flippable");
        for (int l = 0; l < 9; ++l) M.fpar[(((size_t)slot * NQ + q) * 9 + l) * nbp + m] = par[fsp[m] * 9 + l];
      }
    }
    M.fsp[m] = (frozen_sp && frozen_sp[m] >= 0) ? frozen_sp[m] : -1;
    M.fsp[(size_t)nbp + m] = (frozen_sp2 && frozen_sp2[m] >= 0) ? frozen_sp2[m] : -1;
    for (size_t q = 0; q < mt.size(); ++q) M.mtm[q * nbp + m] = mt[q];
    const double* dp = pm + (size_t)nsp * 9;
    for (int s = 1; s < nsp; ++s)
      for (int l = 0; l < 9; ++l)
        if (dp[s * 9 + l] != dp[l]) M.uniform_d = false;
    for (int l = 0; l < 9; ++l) M.dvec[(size_t)l * nbp + m] = dp[l];
    for (int g = 0; g < op->nop; ++g)
      for (int l = 0; l < 9; ++l) M.dperm[((size_t)g * 9 + l) * nbp + m] = dp[op->pi_h[g * 9 + l]];
    for (int q = 0; q < nsp * 9; ++q) {
      if (pm[2 * nsp * 9 + q] != 0.0) M.zero_o = false;
      if (pm[3 * nsp * 9 + q] != 0.0) M.zero_e = false;
    }
    if (z) M.z[m] = cplx{z[2 * m], z[2 * m + 1]};
  }
  return 0;
}

struct MemBuffers {
  uint64_t* chim = nullptr;
  double *wts = nullptr, *fpar = nullptr, *dvec = nullptr, *dperm = nullptr, *mtm = nullptr;
  int* fsp = nullptr;
  cplx* z = nullptr;
  ~MemBuffers() {
    if (mtm) (void)hipFree(mtm);
    if (fsp) (void)hipFree(fsp);
    if (dperm) (void)hipFree(dperm);
    if (chim) (void)hipFree(chim);
    if (wts) (void)hipFree(wts);
    if (fpar) (void)hipFree(fpar);
    if (dvec) (void)hipFree(dvec);
    if (z) (void)hipFree(z);
  }
};

int upload_members(const asr_gpu_op* op, const Members& M, MemBuffers& B, MemDev& md) {
  if (dalloc(&B.chim, M.chim.size()) || dalloc(&B.wts, M.wts.size()) || dalloc(&B.fpar, M.fpar.size()) ||
      dalloc(&B.dvec, M.dvec.size()) || dalloc(&B.z, M.z.size()) ||
      dalloc(&B.dperm, M.dperm.size()) || dalloc(&B.mtm, M.mtm.size()) || dalloc(&B.fsp, M.fsp.size()))
    return 1;
  HCHECK(hipMemcpyAsync(B.fsp, M.fsp.data(), M.fsp.size() * 4, hipMemcpyHostToDevice, op->stream));
  HCHECK(hipMemcpyAsync(B.mtm, M.mtm.data(), M.mtm.size() * 8, hipMemcpyHostToDevice, op->stream));
  HCHECK(hipMemcpyAsync(B.dperm, M.dperm.data(), M.dperm.size() * 8, hipMemcpyHostToDevice, op->stream));
  HCHECK(hipMemcpyAsync(B.chim, M.chim.data(), M.chim.size() * 8, hipMemcpyHostToDevice, op->stream));
  HCHECK(hipMemcpyAsync(B.wts, M.wts.data(), M.wts.size() * 8, hipMemcpyHostToDevice, op->stream));
  HCHECK(hipMemcpyAsync(B.fpar, M.fpar.data(), M.fpar.size() * 8, hipMemcpyHostToDevice, op->stream));
  HCHECK(hipMemcpyAsync(B.dvec, M.dvec.data(), M.dvec.size() * 8, hipMemcpyHostToDevice, op->stream));
  HCHECK(hipMemcpyAsync(B.z, M.z.data(), M.z.size() * sizeof(cplx), hipMemcpyHostToDevice, op->stream));
  md.nb = M.nb; md.nbp = M.nbp; md.mt = M.mt; md.pair = M.pair; md.tact = nullptr; md.ntl = 0;
  md.chimask = B.chim; md.wts = B.wts; md.fpar = B.fpar; md.dvec = B.dvec; md.z = B.z;
  md.dperm = B.dperm;
  md.mtm = B.mtm;
  md.fsp = B.fsp;
  HCHECK(hipStreamSynchronize(op->stream));  // the tables are read from other streams too
  return 0;
}

size_t fused_lds_bytes(const asr_gpu_op* op, size_t tsize) {
  size_t nblk = (size_t)op->ntype * (op->nslot + 1) * 81;
  return nblk * 8 + (size_t)((op->nop * 9 + 1) & ~1) * 4 + (size_t)op->nop * 9 * 8 + BLOCK * tsize;
}

dim3 node_grid(const asr_gpu_op* op, const MemDev& md) {
  int npw = WAVE / md.mt;
  int nodes_per_block = npw * (BLOCK / WAVE);
  return dim3((op->nn + nodes_per_block - 1) / nodes_per_block, md.ntl > 0 ? md.ntl : md.nbp / md.mt, 1);
}

// Optional per-kernel timing (HIP events on the engine's stream), used by asr_gpu_profile_level.
struct KTimer {
  std::vector<hipEvent_t> ev;
  std::vector<int> kind;
  int ord = 0;
};
KTimer* g_timer = nullptr;
std::vector<double> g_acc;  // summed duration per kernel ordinal of a level
std::vector<int> g_cnt;     // levels in which that ordinal ran (the rescale runs every 16 levels only)
int g_levels = 0;

enum { KIND_SPMM = 0, KIND_RAND = 1, KIND_FIN_A = 2, KIND_UPDATE = 3, KIND_FIN_B = 4, KIND_SCALE = 5, KIND_FIX = 6,
       KIND_SPMM_UPD = 7 };

int timer_mark(hipStream_t s, int kind, bool start) {
  if (!g_timer) return 0;
  if (start) {
    size_t need = 2 * (size_t)(g_timer->ord + 1);
    while (g_timer->ev.size() < need) {
      hipEvent_t e;
      HCHECK(hipEventCreate(&e));
      g_timer->ev.push_back(e);
    }
    if (g_timer->kind.size() < (size_t)g_timer->ord + 1) g_timer->kind.push_back(kind);
    HCHECK(hipEventRecord(g_timer->ev[2 * g_timer->ord], s));
  } else {
    HCHECK(hipEventRecord(g_timer->ev[2 * g_timer->ord + 1], s));
    g_timer->ord++;
  }
  return 0;
}

FArgs fargs0() {
  FArgs f;
  std::memset(&f, 0, sizeof(f));
  return f;
}

// A chain vector of the operator's node set (nn + 1 rows of 9 x nbp values of es bytes) is below 4 GiB: the raw buffer
// loads of k_tile and s_links_piid address it with unsigned 32-bit offsets and record counts.  Larger vectors take the
// kernels that index through 64-bit pointers.
bool buffer_ok(const asr_gpu_op* op, const MemDev& md, size_t es) {
  return (size_t)(op->nn + 1) * 9 * md.nbp * es < ((size_t)1 << 32);
}

// the operator pass runs k_tile (member tile 8 or 16, 14 or 18 slots, at most 64 operations, vectors under 4 GiB)
bool tile_kernel(const asr_gpu_op* op, const MemDev& md, size_t es) {
  return !(md.mt == WAVE) && (md.mt == 8 || md.mt == 16) && (op->nslot == 18 || op->nslot == 14) && op->nop <= 64 &&
         buffer_ok(op, md, es);
}

template <class T>
int launch_fused(const asr_gpu_op* op, const MemDev& md, const FArgs& fa, hipStream_t strm) {
  if (fa.out2 && !tile_kernel(op, md, sizeof(T))) return fail("operator pass: out2 needs the tile kernel");
  if (fa.x2 && !std::is_same<T, cplx>::value) return fail("operator pass: energy-linear out2 terms need complex members");
  dim3 grid = node_grid(op, md);
  size_t lds = fused_lds_bytes(op, sizeof(T));
  if ((fa.upd || fa.apart) && !tile_kernel(op, md, sizeof(T))) return fail("operator pass: the fused level needs the tile kernel");
  if (fa.tvv && (!fa.upd || fa.out2 || !fa.partial3 || !std::is_same<T, cplx>::value))
    return fail("operator pass: the block dot needs the fused update of complex members");
  if ((fa.lcn > 0 || fa.out3) && (!std::is_same<T, double>::value || !tile_kernel(op, md, sizeof(T)) || fa.out2 || fa.upd ||
                                   (fa.lcn > 0 && fa.out3) || fa.lcn > 3))
    return fail("operator pass: output combinations need real members, the tile kernel, a plain output and one form");
  if (timer_mark(strm, fa.upd ? KIND_SPMM_UPD : fa.xs ? KIND_SPMM : KIND_RAND, true)) return 1;
  const OpDev od = dev_of(op);
  const double* bl = op->blk;
  const size_t blds_bytes = (size_t)op->ntype * (op->nslot + 1) * 90 * 8;
  // the buffer-load S pass of k_pair and of the real k_node (s_links_piid) needs a vector below 4 GiB; NS = 0 instances
  // take the pointer path
  const bool bok = buffer_ok(op, md, sizeof(T));
  if (md.mt == WAVE && md.pair && sizeof(T) == sizeof(cplx)) {
    grid.y = md.nbp / (WAVE / 2);
#define ASR_PAIR(PI, NS) hipLaunchKernelGGL((k_pair<PI, NS>), grid, dim3(BLOCK), (PI && NS > 0) ? blds_bytes : 0, strm, od, md, \
                                         fa, bl, md.dperm)
    if (op->piid && op->nslot == 18 && bok) ASR_PAIR(true, 18);
    else if (op->piid && op->nslot == 14 && bok) ASR_PAIR(true, 14);
    else if (op->piid) ASR_PAIR(true, 0);
    else if (op->nslot == 18) ASR_PAIR(false, 18);
    else ASR_PAIR(false, 0);
#undef ASR_PAIR
  } else if (md.mt == WAVE) {
#define ASR_NODE(PI, NS) hipLaunchKernelGGL((k_node<T, PI, NS>), grid, dim3(BLOCK), \
                                         (PI && NS > 0 && sizeof(T) == 8) ? blds_bytes : 0, strm, od, md, fa, bl, md.dperm)
    const bool nbok = bok || sizeof(T) != sizeof(double);  // complex members of k_node index through pointers
    if (op->piid && op->nslot == 18 && nbok) ASR_NODE(true, 18);
    else if (op->piid && op->nslot == 14 && nbok) ASR_NODE(true, 14);
    else if (op->piid) ASR_NODE(true, 0);
    else if (op->nslot == 18) ASR_NODE(false, 18);
    else ASR_NODE(false, 0);
#undef ASR_NODE
  } else if (tile_kernel(op, md, sizeof(T))) {
#define ASR_CT4(MT, NS, O2, PI) hipLaunchKernelGGL((k_tile<T, MT, NS, O2, PI>), grid, dim3(BLOCK), \
                                                blds_bytes + (size_t)MT * (2 * op->nop + 1) * 9 * 8, strm, od, md, fa, \
                                                (const double*)op->blkp + (size_t)op->ntype * (op->nslot + 1) * 90)
#define ASR_CT3(MT, NS, O2) do { if (op->piid) ASR_CT4(MT, NS, O2, true); else ASR_CT4(MT, NS, O2, false); } while (0)
#define ASR_CT(MT, NS) do { if constexpr (std::is_same<T, cplx>::value) { if (fa.tvv) { ASR_CT3(MT, NS, 3); break; } } \
                            if (!fa.out2) ASR_CT3(MT, NS, 0); else if (!fa.x2) ASR_CT3(MT, NS, 1); \
                            else ASR_CT3(MT, NS, 2); } while (0)
    if (md.mt == 8 && op->nslot == 18) ASR_CT(8, 18);
    else if (md.mt == 8) ASR_CT(8, 14);
    else if (op->nslot == 18) ASR_CT(16, 18);
    else ASR_CT(16, 14);
#undef ASR_CT
#undef ASR_CT3
#undef ASR_CT4
  } else {
    hipLaunchKernelGGL(k_fused<T>, grid, dim3(BLOCK), lds, strm, dev_of(op), md, fa);
  }
  HCHECK(hipGetLastError());
  if (timer_mark(strm, 0, false)) return 1;
  return 0;
}

// The output combinations of the last pass of H = c + d S d with species-dependent d (real members, o = E_nu = 0):
// the FArgs fields lcn ... rib (see there).
struct RealEpi {
  bool t1_ready = false;  // t1 = rand_d(x~) is given (k_adj_dpass): the first pass is skipped
  int lcn = 0;
  const double* lcv[3] = {nullptr, nullptr, nullptr};
  const double* lcc[3] = {nullptr, nullptr, nullptr};
  double* out3 = nullptr;
  const double *u3 = nullptr, *up3 = nullptr, *ra = nullptr, *rb = nullptr, *rib = nullptr;
};

// the last pass of H x writes y itself (no h1 pass): species-dependent d, no o, E_nu or X(z) term, the tile kernel
bool merged_last_pass(const asr_gpu_op* op, const Members& M, const MemDev& md, size_t es, bool lin) {
  return !M.uniform_d && M.zero_o && M.zero_e && !lin && es == sizeof(double) && tile_kernel(op, md, es);
}

// y = H x (augcls_apply) for every member, normalised coordinates; w_dot: partial <dotu, y>.  Work vectors: t1, t2, h1.
// lin: add X(z) x (complex members, energy-dependent diagonal of route edc).  epi: output combinations of the last pass
// (needs merged_last_pass).
template <class T>
int apply_chain_op(const asr_gpu_op* op, const Members& M, const MemDev& md, const T* x, T* y, T* t1, T* t2, T* h1, bool lin,
                   const T* dotu, T* partial, hipStream_t strm, const RealEpi* epi = nullptr) {
  const bool merged = merged_last_pass(op, M, md, sizeof(T), lin);
  if (epi && !merged) return fail("operator pass: output combinations need the merged last pass");
  // h1 = h x~ = rand_c(x~) + D S D x~; with the tile kernel the same launch writes t1 = rand_o(h1) at the nodes without
  // flip partners (99 % of them), and k_rand_nodes the rest, in place of a full rand_o pass
  bool t1_done = false;
  if (M.uniform_d) {
    FArgs f = fargs0();
    f.xs = x; f.xs_norm = 1; f.s_din = 1; f.s_dout = 1; f.s_coef = 1.0;
    f.xr[0] = x; f.qr[0] = QC; f.r_norm[0] = 1; f.r_coef[0] = 1.0; f.nr = 1;
    f.out = h1;
    if (!M.zero_o && tile_kernel(op, md, sizeof(T))) {
      f.out2 = t1; f.q2 = QO; f.c2 = 1.0;
      t1_done = true;
    }
    if (launch_fused<T>(op, md, f, strm)) return 1;
    if (t1_done && op->nplist > 0) {
      if (timer_mark(strm, KIND_FIX, true)) return 1;
      const size_t nt = (size_t)op->nplist * md.nbp;
      hipLaunchKernelGGL(k_rand_nodes<T>, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, strm, dev_of(op), md,
                         (const int*)op->plist, op->nplist, (const T*)h1, t1, (int)QO, 1.0, (const T*)nullptr, 0, 0, 0.0);
      HCHECK(hipGetLastError());
      if (timer_mark(strm, 0, false)) return 1;
    }
  } else {
    FArgs f = fargs0();
    f.xr[0] = x; f.qr[0] = QD; f.r_norm[0] = 1; f.r_coef[0] = 1.0; f.nr = 1; f.out = t1;
    if (!(epi && epi->t1_ready) && launch_fused<T>(op, md, f, strm)) return 1;
    f = fargs0();
    f.xs = t1; f.s_coef = 1.0; f.out = t2;
    if (launch_fused<T>(op, md, f, strm)) return 1;
    f = fargs0();
    f.xr[0] = x; f.qr[0] = QC; f.r_norm[0] = 1; f.r_coef[0] = 1.0;
    f.xr[1] = t2; f.qr[1] = QD; f.r_norm[1] = 0; f.r_coef[1] = 1.0; f.nr = 2;
    if (merged) {
      // y = sq (rand_c(x~) + rand_d(t2)) at once: with no o and E_nu term the last pass would only scale h1 by sq (the
      // values are bit for bit those of the two passes)
      f.out = y;
      f.out_sq = 1;
      f.dotu = dotu; f.partial = partial;
      if (epi) {
        f.lcn = epi->lcn;
        for (int t = 0; t < 3; ++t) { f.lcv[t] = epi->lcv[t]; f.lcc[t] = epi->lcc[t]; }
        f.out3 = epi->out3; f.u3 = epi->u3; f.up3 = epi->up3; f.ra = epi->ra; f.rb = epi->rb; f.rib = epi->rib;
      }
      return launch_fused<T>(op, md, f, strm);
    }
    f.out = h1;
    if (launch_fused<T>(op, md, f, strm)) return 1;
  }
  // final: y = sq (rand_e(x~) + h1 - h(rand_o(h1)) [+ X(z) x~])
  FArgs f = fargs0();
  int nr = 0;
  if (!M.zero_o) {
    FArgs g = fargs0();
    g.xr[0] = h1; g.qr[0] = QO; g.r_norm[0] = 0; g.r_coef[0] = 1.0; g.nr = 1; g.out = t1;  // r = rand_o(h1)
    if (!t1_done && launch_fused<T>(op, md, g, strm)) return 1;
    if (M.uniform_d) {
      f.xs = t1; f.s_din = 1; f.s_dout = 1; f.s_coef = -1.0;
    } else {
      g = fargs0();
      g.xr[0] = t1; g.qr[0] = QD; g.r_coef[0] = 1.0; g.nr = 1; g.out = t2;  // t2 = rand_d(r)
      if (launch_fused<T>(op, md, g, strm)) return 1;
      g = fargs0();
      g.xs = t2; g.s_coef = 1.0; g.out = y;  // y used as scratch: S rand_d(r)
      if (launch_fused<T>(op, md, g, strm)) return 1;
      f.xr[nr] = y; f.qr[nr] = QD; f.r_norm[nr] = 0; f.r_coef[nr] = -1.0; ++nr;
      // the final kernel reads y at flip partners and writes y: use t2 as its output and copy back below
    }
    f.xr[nr] = t1; f.qr[nr] = QC; f.r_norm[nr] = 0; f.r_coef[nr] = -1.0; ++nr;
  }
  if (!M.zero_e) { f.xr[nr] = x; f.qr[nr] = QE; f.r_norm[nr] = 1; f.r_coef[nr] = 1.0; ++nr; }
  f.nr = nr;
  if (lin) { f.xl[0] = x; f.lq[0] = QPA; f.l_norm[0] = 1; f.l_coef[0] = 1.0; f.nl = 1; }
  f.xo[0] = h1; f.o_coef[0] = 1.0; f.no = 1;
  f.out_sq = 1;
  f.dotu = dotu; f.partial = partial;
  if (!M.zero_o && !M.uniform_d) {
    f.out = t2;
    if (launch_fused<T>(op, md, f, strm)) return 1;
    HCHECK(hipMemcpyAsync(y, t2, sizeof(T) * (size_t)op->nn * 9 * md.nbp, hipMemcpyDeviceToDevice, strm));
  } else {
    f.out = y;
    if (launch_fused<T>(op, md, f, strm)) return 1;
  }
  return 0;
}

// w = z x - K'(z) x for the chains of route ed3 (ed3_kmul and the chain's w = z u - K' u, src/asr_ed3.f90:99-129,169-176):
//   K' x = et x - h'x + u h'x + h'(u x - Qt h'x),  so  w = sq [ (z - et) x~ + h1 - u h1 - h'(t) ],
//   h1 = h' x~ = g x~ + D_c S D_c x~,   t = u x~ - Qt h1,
// in normalised coordinates (x~ = x / sq, h1 and t in amplitude coordinates).  The energy-linear diagonals are the slots
// (QPA, QPB) = z - et, (QUA, QUB) = u, (QQA, QQB) = Qt.  Three launches: the S pass for h1, the diagonal pass for t, the S
// pass on t with the other terms and the dot <x, w>.  h' has one d for every species (D_c^1/2).
template <class T>
int apply_ed3_op(const asr_gpu_op* op, const Members& M, const MemDev& md, const T* x, T* y, T* t1, T* h1, const T* dotu,
                 T* partial, hipStream_t strm) {
  if constexpr (!std::is_same<T, cplx>::value) {
    return fail("route ed3: complex members only");
  } else {
    if (!M.uniform_d) return fail("route ed3: d must be the same for every species (D_c^1/2)");
    FArgs f = fargs0();
    f.xs = x; f.xs_norm = 1; f.s_din = 1; f.s_dout = 1; f.s_coef = 1.0;
    f.xr[0] = x; f.qr[0] = QC; f.r_norm[0] = 1; f.r_coef[0] = 1.0; f.nr = 1;
    f.out = h1;
    // t1 = u x~ - Qt h1: node-local at the nodes without flip partners, so the first S pass writes it there from the
    // registers that hold h1 (and the own row of x) and k_rand_nodes writes the listed nodes
    const bool fuse = tile_kernel(op, md, sizeof(T));
    if (fuse) {
      f.out2 = t1; f.q2 = QQA; f.c2 = -1.0;
      f.x2 = x; f.q2x = QUA; f.x2norm = 1; f.c2x = 1.0;
    }
    if (launch_fused<T>(op, md, f, strm)) return 1;
    if (fuse && op->nplist > 0) {
      if (timer_mark(strm, KIND_FIX, true)) return 1;
      const size_t nt = (size_t)op->nplist * md.nbp;
      hipLaunchKernelGGL(k_rand_nodes<T>, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, strm, dev_of(op), md,
                         (const int*)op->plist, op->nplist, (const T*)h1, t1, (int)QQA, -1.0, x, (int)QUA, 1, 1.0);
      HCHECK(hipGetLastError());
      if (timer_mark(strm, 0, false)) return 1;
    }
    if (!fuse) {
      f = fargs0();
      f.xl[0] = x; f.lq[0] = QUA; f.l_norm[0] = 1; f.l_coef[0] = 1.0;
      f.xl[1] = h1; f.lq[1] = QQA; f.l_norm[1] = 0; f.l_coef[1] = -1.0; f.nl = 2;
      f.out = t1;
      if (launch_fused<T>(op, md, f, strm)) return 1;
    }
    f = fargs0();
    f.xs = t1; f.xs_norm = 0; f.s_din = 1; f.s_dout = 1; f.s_coef = -1.0;
    f.xr[0] = t1; f.qr[0] = QC; f.r_norm[0] = 0; f.r_coef[0] = -1.0; f.nr = 1;
    f.xl[0] = x; f.lq[0] = QPA; f.l_norm[0] = 1; f.l_coef[0] = 1.0;
    f.xl[1] = h1; f.lq[1] = QUA; f.l_norm[1] = 0; f.l_coef[1] = -1.0; f.nl = 2;
    f.xo[0] = h1; f.o_coef[0] = 1.0; f.no = 1;
    f.out_sq = 1;
    f.dotu = dotu; f.partial = partial;
    f.out = y;
    return launch_fused<T>(op, md, f, strm);
  }
}

// the operator of one chain level: route ed3's z - K'(z), or H (+ X(z)) of the real and edc chains
template <class T>
int apply_level_op(const asr_gpu_op* op, const Members& M, const MemDev& md, const T* x, T* y, T* t1, T* t2, T* h1, bool lin,
                   const T* dotu, T* partial, hipStream_t strm) {
  if (M.ed3) return apply_ed3_op<T>(op, M, md, x, y, t1, h1, dotu, partial, strm);
  return apply_chain_op<T>(op, M, md, x, y, t1, t2, h1, lin, dotu, partial, strm);
}

// The autotuner of the batched chains (the role of chain_teams / blk_nbmax, src/asr_chain.f90:140-155,
// src/asr_augblk.f90:32-36): the member tile MT (members per wave, 64 / MT nodes per wave) sets the cache footprint of a
// node (9 MT values) against the lane occupancy of ragged member counts.  Every candidate runs one operator application
// of the level (same kernels, same members) between HIP events; the fastest is kept per (operator, member count, kind).
template <class T>
int autotune_tile(asr_gpu_op* op, int nchain, const int* start_orb, const double* params, const double* pab,
                  const int* frozen_sp, const double* conc, const double* z, int* best, const double* pex = nullptr,
                  const int* fsp2 = nullptr, const uint64_t* chim = nullptr) {
  Members M0;
  if (prepare_members(op, nchain, start_orb, params, pab, frozen_sp, conc, z, M0, 0, pex, fsp2, chim)) return 1;
  long long key = ((long long)nchain << 8) | (sizeof(T) == 16 ? 1 : 0) | (M0.uniform_d ? 2 : 0) | (M0.zero_o ? 4 : 0) |
                  (pab ? 8 : 0) | (pex ? 16 : 0);
  for (auto& kv : op->tuned)
    if (kv.first == key) { *best = kv.second; return 0; }
  const int top = pow2_tile(nchain);
  double best_ms = 1e300;
  int best_mt = top;
  std::vector<int> cands;
  for (int mt = std::min(8, top); mt <= top; mt <<= 1) cands.push_back(mt);
  if (top == WAVE && sizeof(T) == sizeof(cplx)) cands.push_back(65);
  for (int mt : cands) {
    Members M;
    if (prepare_members(op, nchain, start_orb, params, pab, frozen_sp, conc, z, M, mt, pex, fsp2, chim)) return 1;
    MemBuffers MB;
    MemDev md;
    if (upload_members(op, M, MB, md)) return 1;
    const size_t nv = (size_t)(op->nn + 1) * 9 * M.nbp;  // + the zero padding node
    T *x = nullptr, *y = nullptr, *t1 = nullptr, *t2 = nullptr, *h1 = nullptr;
    BlockPool* pool = &op->pool;
    if (palloc(pool, &x, nv) || palloc(pool, &y, nv) || palloc(pool, &t1, nv) || palloc(pool, &t2, nv) ||
        palloc(pool, &h1, nv))
      return 1;
    int st = 0;
    hipEvent_t e0, e1;
    st |= hipEventCreate(&e0) != hipSuccess;
    st |= hipEventCreate(&e1) != hipSuccess;
    for (T* p : {x, y, t1, t2, h1}) st |= hipMemsetAsync(p, 0, nv * sizeof(T), op->stream) != hipSuccess;
    if (!st) st = apply_level_op<T>(op, M, md, x, y, t1, t2, h1, pab != nullptr, nullptr, nullptr, op->stream);
    if (!st) st |= hipEventRecord(e0, op->stream) != hipSuccess;
    for (int r = 0; r < 2 && !st; ++r) st = apply_level_op<T>(op, M, md, x, y, t1, t2, h1, pab != nullptr, nullptr, nullptr, op->stream);
    if (!st) st |= hipEventRecord(e1, op->stream) != hipSuccess;
    if (!st) st |= hipEventSynchronize(e1) != hipSuccess;
    float ms = 0.f;
    if (!st) st |= hipEventElapsedTime(&ms, e0, e1) != hipSuccess;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(op->stream);
    for (T* p : {x, y, t1, t2, h1}) pool->put(p);
    if (st) return fail("autotune: a timing launch failed");
    // a tile that pads many idle members is charged for them
    if (ms < best_ms) { best_ms = ms; best_mt = mt; }
  }
  op->tuned.push_back({key, best_mt});
  *best = best_mt;
  return 0;
}

// status of a chain run that must be repeated with the separate update pass (ChainRun::finish)
constexpr int RERUN_ASYM = 2;

template <class T>
struct ChainBuffers {
  T *u = nullptr, *up = nullptr, *w = nullptr, *t1 = nullptr, *t2 = nullptr, *h1 = nullptr, *part = nullptr;
  T *a = nullptr, *b = nullptr, *a_eff = nullptr, *b_eff = nullptr, *inv_b = nullptr;
  T *tv = nullptr, *part2 = nullptr, *c = nullptr;  // route ed3: projection vector, its dot partials, c_k
  T* lpart = nullptr;  // the fused edc level: a_k terms of the nodes with flip partners (nplist x nbp)
  T* pchk = nullptr;   // ... the partials of the direct <u, H u> of its check
  int* asym = nullptr; // ... set on the device when the check fails
  double *qv = nullptr, *qpart = nullptr, *qcoef = nullptr;  // real chains with full reorthogonalisation
  double* tw = nullptr;  // projected chains: dressing weights ntarget x nsmax x 9 x nbp
  T* cpr = nullptr;      // projected chains: projections per level, ld x ntarget x 9 x nbp
  int *level = nullptr, *depth = nullptr, *active = nullptr, *mused = nullptr, *closed = nullptr;
  BlockPool* pool = nullptr;  // the level vectors u .. h1, tv come from the operator's cache and go back there
  ~ChainBuffers() {
    for (void* p : {(void*)u, (void*)up, (void*)w, (void*)t1, (void*)t2, (void*)h1, (void*)tv})
      if (p) {
        if (pool) pool->put(p);
        else (void)hipFree(p);
      }
    for (void* p : {(void*)part, (void*)a, (void*)b, (void*)a_eff, (void*)b_eff, (void*)inv_b, (void*)level, (void*)depth,
                    (void*)active, (void*)mused, (void*)part2, (void*)c, (void*)closed, (void*)lpart, (void*)pchk,
                    (void*)asym, (void*)qv, (void*)qpart,
                    (void*)qcoef, (void*)tw, (void*)cpr})
      if (p) (void)hipFree(p);
  }
};

// The two-chain blocks of route ed3 (ed3_blocks, src/asr_ed3.f90:247-278) and of the EPI pairs (epi_blocks,
// src/asr_epi.f90:628-647), per member: member 2j is the chain of block j from |a> = |s0> + o_c |v> (normalised), projected
// on |v> = h'|s0>; member 2j + 1 the chain from |v> / |v|, projected on |a>.  The operator is K'(z) of route ed3 when pex
// is given, else H(z) of route edc (o_c = 0 there).
struct BlockIn {
  const double* pex;   // nmember x 4 x nsp x 9: ua, ub, qa, qb (route ed3) or null
  const double* oc;    // nmember
  const int* side;     // nmember: 0 from |a>, 1 from |v>
  const double* v9;    // nmember x 9: the start vector on the start node, or null for the unit vector of the start orbital
  const uint64_t* chim;  // nmember: characters of v9 (bit g set <=> chi(g) = -1) or null (those of the start orbital)
  const int* fsp2;     // nmember: species of the frozen partner site (fz slot 2) or null
};

// One batch of chains on one operator: members sorted by depth (deepest first, so member tiles retire in order), device
// state, and the launches of one level on the group's stream.  Several groups (one per operator) run concurrently on
// their own streams; the whole run can be captured into one HIP graph (all initialisation happens before the capture).
template <class T>
struct ChainRun {
  asr_gpu_op* op = nullptr;
  hipStream_t stream = nullptr;
  Members M;
  MemBuffers MB;
  MemDev md;
  ChainBuffers<T> B;
  ChainDev cd;
  dim3 grid, sgrid;
  int nbx = 0, apply_mt = 1, maxdepth = 0, ld = 0, project = 0, nchain = 0, ntile = 1;
  size_t nv = 0;
  bool lin = false;
  std::vector<int> order;  // sorted member -> input chain
  int* tact = nullptr;     // per member tile: 1 while any member of the tile is active
  int level_host = 0;      // levels issued
  std::vector<cplx> gprev, fprev;  // adaptive depth: G(z) (and the projection f of the blocks) at the previous tested level
  // the adaptive depth on the device (tails hull, avg2, none): no host synchronisation, so the run can be one HIP graph
  bool dev_depth = false;
  cplx *gprev_d = nullptr, *fprev_d = nullptr;
  double* tnorm_d = nullptr;
  int* conv_d = nullptr;                 // members the last test found converged (they leave at the next k_fin_a, or
                                         // in a window run at the refill after the test's level)
  hipStream_t tstream = nullptr;         // the test's own stream (default priority)
  hipEvent_t ev_a = nullptr, ev_t = nullptr;
  bool test_pending = false;
  bool ed3 = false;                  // two-chain blocks (start vectors, projections c_k, closure flags)
  std::vector<double> snorm, tnorm;  // blocks: norms of the start vector and of the projection vector per member
  std::vector<double> s0h;           // blocks: 9 x nbp start vectors on the start node
  bool reorth = false;               // real chains: full reorthogonalisation against every stored Lanczos vector
  size_t qplane = 0;
  bool proj = false;                 // projected chains: projections of every level's vector on the target states
  // lagged normalisation of the complex chains: the vector of level k is kept as s_k u_k (w' of the previous update, not
  // divided by b), the per-member scale s_k = s_{k-1} b_{k-1} enters only the scalar bookkeeping (a_k, b_k, the update's
  // coefficients, c_k, the projections); every 16 levels the vector is normalised once (k_rescale).  The scales rotate
  // over three buffers: sc3[k % 3] = s_k, sc3[(k + 2) % 3] = s_{k-1}, sc3[(k + 1) % 3] receives s_{k+1}.
  bool lag = false;
  // The fused level (complex edc chains on the member-tile pass): a_k = <x, H x> / s_k^2 from the first operator pass, by
  // the symmetry of Sq^2 h and Sq^2 X_o on the symmetric subspace,
  //   <x~, Sq^2 (E + X(z) + h - h X_o h) x~> = sum_i sq_i^2 [x~ (E + X(z)) x~ + x~ h1 - h1 (X_o h1)]_i,  h1 = h x~,
  // so the second pass applies the three-term update, the projection and the partials of <w, w> to its output in
  // registers, and the separate update pass (a read of w, u, u_prev and a write of w per level) goes away.
  bool fused = false;
  T* sc3[3] = {nullptr, nullptr, nullptr};
  const T* s_cur() const { return lag ? sc3[level_host % 3] : nullptr; }
  const T* s_prev() const { return lag ? sc3[(level_host + 2) % 3] : nullptr; }
  T* s_next() const { return lag ? sc3[(level_host + 1) % 3] : nullptr; }

  // member compaction (host-issued levels with the adaptive depth): the permutation of the last compaction and a staging
  // buffer for the in-place permutations of the member tables
  int* perm_d = nullptr;
  void* ptmp = nullptr;
  size_t ptmp_bytes = 0;
  int ncompact = 0;
  // window run (route edc with the adaptive depth on the device, more chains than op->window): the run holds
  // op->window member slots; the chains wait in a queue (sorted order) and take the slots of members that have left, whose
  // results go to the rows of their chains (a_all, b_all, mused_all, relative levels); a slot's member started at level
  // slot_s0 (s0_d on the device), and the tables of every chain stay on the device (Mall / MBall / mdall)
  bool win = false;
  Members Mall;
  MemBuffers MBall;
  MemDev mdall{};
  std::vector<int> so_all, nl_all, slot_chain, slot_s0, slot_dirty;
  int qnext = 0;
  T *a_all = nullptr, *b_all = nullptr;
  int *mused_all = nullptr, *s0_d = nullptr, *wslots = nullptr, *wchains = nullptr, *wvals = nullptr, *wl0 = nullptr;
  // the queue on the device (op->queue_dev): k_refill after every level while chains wait, the host reads its counters
  // every 8 levels; once the queue is empty the run continues on the host path (refill, then compaction).  wq_d: [0] next
  // chain, [1] extractions, [2] dirty starts, [3] starts, [4] live members, [5] 1 while anything remains
  bool dev_queue = false;
  int *wq_d = nullptr, *slot_chain_d = nullptr, *slot_dirty_d = nullptr, *nl_all_d = nullptr, *so_all_d = nullptr,
      *xslots_d = nullptr, *xchains_d = nullptr;

  ~ChainRun() {
    // work of an aborted run may still be queued: the cached blocks go back only once the stream is idle
    if (stream) (void)hipStreamSynchronize(stream);
    if (tact) (void)hipFree(tact);
    if (perm_d) (void)hipFree(perm_d);
    if (ptmp) (void)hipFree(ptmp);
    if (tstream) (void)hipStreamSynchronize(tstream);
    if (conv_d) (void)hipFree(conv_d);
    if (ev_a) (void)hipEventDestroy(ev_a);
    if (ev_t) (void)hipEventDestroy(ev_t);
    if (tstream) (void)hipStreamDestroy(tstream);
    for (void* p : {(void*)a_all, (void*)b_all, (void*)mused_all, (void*)s0_d, (void*)wslots, (void*)wchains, (void*)wvals,
                    (void*)wl0, (void*)wq_d, (void*)slot_chain_d, (void*)slot_dirty_d, (void*)nl_all_d, (void*)so_all_d,
                    (void*)xslots_d, (void*)xchains_d})
      if (p) (void)hipFree(p);
    for (void* p : {(void*)gprev_d, (void*)fprev_d, (void*)tnorm_d})
      if (p) (void)hipFree(p);
    for (T* p : sc3)
      if (p) (void)hipFree(p);
  }

  int setup(asr_gpu_op* op_, hipStream_t s, int nch, const int* start_orb, const double* params, const double* pab,
            const int* frozen_sp, const double* conc, const double* z, const int* nlev, int project_, int ldab, int tile,
            double tiny, bool real_chain, const BlockIn* e3 = nullptr, const double* tw_in = nullptr,
            bool allow_window = false) {
    op = op_; stream = s; nchain = nch; project = project_; ld = ldab; lin = (pab != nullptr);
    op->ncompact = 0;
    proj = (tw_in != nullptr);
    if (proj && (op->ntarget <= 0 || !std::is_same<T, cplx>::value)) return fail("projected chains: no target states set");
    ed3 = (e3 != nullptr);
    if (ed3 && (!std::is_same<T, cplx>::value || !pab || !z)) return fail("chain blocks: complex chains with X(z) and z");
    if (!op->have_tables || !op->have_ops || !op->have_blocks || !op->have_classes)
      return fail("asr_gpu chains: operator tables incomplete");
    // deepest chains first
    order.resize(nch);
    for (int c = 0; c < nch; ++c) order[c] = c;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return nlev[x] > nlev[y]; });
    const int nsp9 = op->nsp * 9, ncn = op->ncls * op->nsmax;
    std::vector<int> so(nch), fz(nch), nl(nch);
    std::vector<double> pp((size_t)nch * 4 * nsp9), pa(pab ? (size_t)nch * 2 * nsp9 : 0), cc(conc ? (size_t)nch * ncn : 0),
        zz(z ? 2 * nch : 0), px(ed3 && e3->pex ? (size_t)nch * 4 * nsp9 : 0), ocs(ed3 ? nch : 0),
        v9s(ed3 && e3->v9 ? (size_t)nch * 9 : 0);
    std::vector<int> sides(ed3 ? nch : 0), fz2(ed3 && e3->fsp2 ? nch : 0);
    std::vector<uint64_t> chs(ed3 && e3->chim ? nch : 0);
    for (int q = 0; q < nch; ++q) {
      const int c = order[q];
      so[q] = start_orb[c];
      fz[q] = frozen_sp ? frozen_sp[c] : -1;
      nl[q] = nlev[c];
      std::memcpy(&pp[(size_t)q * 4 * nsp9], params + (size_t)c * 4 * nsp9, sizeof(double) * 4 * nsp9);
      if (pab) std::memcpy(&pa[(size_t)q * 2 * nsp9], pab + (size_t)c * 2 * nsp9, sizeof(double) * 2 * nsp9);
      if (conc) std::memcpy(&cc[(size_t)q * ncn], conc + (size_t)c * ncn, sizeof(double) * ncn);
      if (z) { zz[2 * q] = z[2 * c]; zz[2 * q + 1] = z[2 * c + 1]; }
      if (ed3) {
        if (e3->pex) std::memcpy(&px[(size_t)q * 4 * nsp9], e3->pex + (size_t)c * 4 * nsp9, sizeof(double) * 4 * nsp9);
        if (e3->v9) std::memcpy(&v9s[(size_t)q * 9], e3->v9 + (size_t)c * 9, sizeof(double) * 9);
        if (e3->chim) chs[q] = e3->chim[c];
        if (e3->fsp2) fz2[q] = e3->fsp2[c];
        ocs[q] = e3->oc[c];
        sides[q] = e3->side[c];
      }
    }
    const double* pab_s = pab ? pa.data() : nullptr;
    const double* conc_s = conc ? cc.data() : nullptr;
    const double* z_s = z ? zz.data() : nullptr;
    const double* px_s = px.empty() ? nullptr : px.data();
    const int* fz2_s = fz2.empty() ? nullptr : fz2.data();
    const uint64_t* ch_s = chs.empty() ? nullptr : chs.data();
    maxdepth = 0;
    for (int q = 0; q < nch; ++q) maxdepth = std::max(maxdepth, nl[q]);
    if (maxdepth > ldab) return fail("chains: depth larger than the leading dimension of a, b");
    // a window run: the route-edc chains with the adaptive depth on the device hold op->window slots, the others queue
    win = allow_window && std::is_same<T, cplx>::value && op->window > 0 && nch > op->window && lin && !proj && !ed3 &&
          !real_chain && op->depth_tol > 0.0 && op->depth_tail != 1 && maxdepth <= DEPTH_MAX;
    const int nmem = win ? op->window : nch;  // members held at once
    if (tile == 0 && autotune_tile<T>(op, nmem, so.data(), pp.data(), pab_s, fz.data(), conc_s, z_s, &tile, px_s, fz2_s, ch_s))
      return 1;
    if (!win) {
      if (prepare_members(op, nch, so.data(), pp.data(), pab_s, fz.data(), conc_s, z_s, M, tile, px_s, fz2_s, ch_s)) return 1;
    } else {
      // the tables of every chain stay on the device (stride Mall.nbp); the slots take the first nmem chains
      if (prepare_members(op, nch, so.data(), pp.data(), pab_s, fz.data(), conc_s, z_s, Mall, tile, px_s, fz2_s, ch_s))
        return 1;
      if (prepare_members(op, nmem, so.data(), pp.data(), pab_s, fz.data(), conc_s, z_s, M, tile, px_s, fz2_s, ch_s))
        return 1;
      if (M.mt != Mall.mt) return fail("window run: member tile mismatch");
      if (upload_members(op, Mall, MBall, mdall)) return 1;
      so_all = so;
      nl_all = nl;
      slot_chain.assign(M.nbp, -1);
      slot_s0.assign(M.nbp, 0);
      slot_dirty.assign(M.nbp, 0);
      for (int q = 0; q < nmem; ++q) slot_chain[q] = q;
      qnext = nmem;
      if (dalloc(&a_all, (size_t)nch * ldab) || dalloc(&b_all, (size_t)nch * ldab) || dalloc(&mused_all, nch) ||
          dalloc(&s0_d, M.nbp) || dalloc(&wslots, M.nbp) || dalloc(&wchains, M.nbp) || dalloc(&wvals, M.nbp) ||
          dalloc(&wl0, M.nbp))
        return 1;
      HCHECK(hipMemsetAsync(a_all, 0, (size_t)nch * ldab * sizeof(T), stream));
      HCHECK(hipMemsetAsync(b_all, 0, (size_t)nch * ldab * sizeof(T), stream));
      HCHECK(hipMemsetAsync(mused_all, 0, nch * sizeof(int), stream));
      HCHECK(hipMemsetAsync(s0_d, 0, M.nbp * sizeof(int), stream));
    }
    if (upload_members(op, M, MB, md)) return 1;
    nv = (size_t)(op->nn + 1) * 9 * M.nbp;  // + the zero padding node (target of missing links)
    grid = node_grid(op, md);
    nbx = grid.x;
    ntile = M.nbp / M.mt;
    const size_t npart = (size_t)grid.x * grid.y * md.mt;
    // the dot partials of the apply pass: complex members with tile 64 run as lane pairs, 32 members per wave
    apply_mt = (md.mt == WAVE && md.pair && sizeof(T) == sizeof(cplx)) ? WAVE / 2 : md.mt;
    B.pool = &op->pool;
    if (palloc(B.pool, &B.u, nv) || palloc(B.pool, &B.up, nv) || palloc(B.pool, &B.w, nv) || palloc(B.pool, &B.t1, nv) ||
        palloc(B.pool, &B.t2, nv) || palloc(B.pool, &B.h1, nv) || dalloc(&B.part, npart) || dalloc(&B.a, (size_t)M.nbp * ldab) ||
        dalloc(&B.b, (size_t)M.nbp * ldab) || dalloc(&B.a_eff, M.nbp) || dalloc(&B.b_eff, M.nbp) ||
        dalloc(&B.inv_b, M.nbp) || dalloc(&B.level, 1) || dalloc(&B.depth, M.nbp) || dalloc(&B.active, M.nbp) ||
        dalloc(&B.mused, M.nbp) || dalloc(&tact, std::max(ntile, 2)))
      return 1;
    if (ed3 && (palloc(B.pool, &B.tv, nv) || dalloc(&B.part2, npart) || dalloc(&B.c, (size_t)M.nbp * ldab) ||
                dalloc(&B.closed, M.nbp)))
      return 1;
    if (win) {
      // the chains that end early (b_k < tiny) are flagged: their slot's vector is not zero when the next chain takes it
      if (dalloc(&B.closed, M.nbp)) return 1;
      HCHECK(hipMemsetAsync(B.closed, 0, M.nbp * sizeof(int), stream));
    }
    reorth = op->reorth && real_chain && std::is_same<T, double>::value;
    if (reorth) {
      // the stored Lanczos vectors of every member: maxdepth x nv doubles (lanczos keeps q(dim, nmax) per chain)
      if ((double)maxdepth * nv * 8.0 > 64.0 * (1ull << 30))
        return fail("lanczos reorth: the stored vectors would exceed 64 GiB (full reorthogonalisation is for small cases)");
      qplane = npart;
      if (dalloc(&B.qv, (size_t)maxdepth * nv) || dalloc(&B.qpart, (size_t)maxdepth * npart) ||
          dalloc(&B.qcoef, (size_t)maxdepth * M.nbp))
        return 1;
    }
    if (proj) {
      // dressing weights of the sorted members, [t][ks][l][member]; projections [level][t][l][member]
      const int nt = op->ntarget, nsm = op->nsmax;
      std::vector<double> twd((size_t)nt * nsm * 9 * M.nbp, 0.0);
      for (int q = 0; q < nch; ++q) {
        const double* src = tw_in + (size_t)order[q] * nt * nsm * 9;
        for (size_t e = 0; e < (size_t)nt * nsm * 9; ++e) twd[e * M.nbp + q] = src[e];
      }
      if (dalloc(&B.tw, twd.size()) || dalloc(&B.cpr, (size_t)ldab * nt * 9 * M.nbp)) return 1;
      HCHECK(hipMemcpyAsync(B.tw, twd.data(), twd.size() * 8, hipMemcpyHostToDevice, stream));
      HCHECK(hipMemsetAsync(B.cpr, 0, (size_t)ldab * nt * 9 * M.nbp * sizeof(T), stream));
    }
    lag = std::is_same<T, cplx>::value && !real_chain;
    // the fused level of the edc chains and of the chain blocks (route ed3, EPI): the member-tile operator pass with one
    // d for every species
    fused = lag && lin && project && !reorth && M.uniform_d && op->level_form == 0 && !op->asym &&
            tile_kernel(op, md, sizeof(T));
    if (fused) {
      if (op->nplist > 0 && dalloc(&B.lpart, (size_t)op->nplist * M.nbp)) return 1;
      if (dalloc(&B.pchk, npart) || dalloc(&B.asym, 1)) return 1;
      HCHECK(hipMemsetAsync(B.asym, 0, sizeof(int), stream));
    }
    std::vector<T> one;  // lives until the stream synchronisation below (source of asynchronous copies)
    if (lag) {
      one.resize(M.nbp);
      std::memset(one.data(), 0, one.size() * sizeof(T));
      for (int q = 0; q < M.nbp; ++q) *(double*)&one[q] = 1.0;
      for (T*& p : sc3) {
        if (dalloc(&p, M.nbp)) return 1;
        HCHECK(hipMemcpyAsync(p, one.data(), M.nbp * sizeof(T), hipMemcpyHostToDevice, stream));
      }
    }
    // u (the start vectors) and u_prev (read by the first update with b = 0) start at zero; the level's outputs w, t1, t2,
    // h1 are written at every node of a running tile before they are read, and only their padding node's row, which the
    // gathers read for missing links, must be zero
    for (T* q : {B.u, B.up}) HCHECK(hipMemsetAsync(q, 0, nv * sizeof(T), stream));
    for (T* q : {B.w, B.t1, B.t2, B.h1})
      HCHECK(hipMemsetAsync(q + (size_t)op->nn * 9 * M.nbp, 0, (size_t)9 * M.nbp * sizeof(T), stream));
    HCHECK(hipMemsetAsync(B.a, 0, (size_t)M.nbp * ldab * sizeof(T), stream));
    HCHECK(hipMemsetAsync(B.b, 0, (size_t)M.nbp * ldab * sizeof(T), stream));
    for (T* q : {B.a_eff, B.b_eff, B.inv_b}) HCHECK(hipMemsetAsync(q, 0, M.nbp * sizeof(T), stream));
    HCHECK(hipMemsetAsync(B.level, 0, sizeof(int), stream));
    HCHECK(hipMemsetAsync(B.mused, 0, M.nbp * sizeof(int), stream));
    std::vector<int> dep(M.nbp, 0), act(M.nbp, 0), ta(std::max(ntile, 2), 0);
    std::vector<T> ustart((size_t)9 * M.nbp);
    std::memset(ustart.data(), 0, ustart.size() * sizeof(T));
    for (int q = 0; q < M.nb; ++q) {
      dep[q] = nl[q];
      act[q] = nl[q] > 0 ? 1 : 0;
      if (act[q]) ta[q / M.mt] = 1;
      // start vector: unit at component iv = 9 * node 0 + l0, or the block's combination on the start node
      if (ed3 && !v9s.empty())
        for (int l = 0; l < 9; ++l) *(double*)&ustart[(size_t)l * M.nbp + q] = v9s[(size_t)q * 9 + l];
      else
        *(double*)&ustart[(size_t)so[q] * M.nbp + q] = 1.0;
    }
    if (ed3) {
      s0h.assign((size_t)9 * M.nbp, 0.0);
      for (int q = 0; q < nch; ++q)
        for (int l = 0; l < 9; ++l) s0h[(size_t)l * M.nbp + q] = *(const double*)&ustart[(size_t)l * M.nbp + q];
    }
    HCHECK(hipMemcpyAsync(B.u, ustart.data(), ustart.size() * sizeof(T), hipMemcpyHostToDevice, stream));
    HCHECK(hipMemcpyAsync(B.depth, dep.data(), M.nbp * sizeof(int), hipMemcpyHostToDevice, stream));
    HCHECK(hipMemcpyAsync(B.active, act.data(), M.nbp * sizeof(int), hipMemcpyHostToDevice, stream));
    HCHECK(hipMemcpyAsync(tact, ta.data(), ta.size() * sizeof(int), hipMemcpyHostToDevice, stream));
    HCHECK(hipStreamSynchronize(stream));
    cd.level = B.level; cd.depth = B.depth; cd.active = B.active; cd.mused = B.mused; cd.ld = ldab;
    cd.real_chain = real_chain ? 1 : 0; cd.tiny = tiny;
    cd.cvec = B.c; cd.closed = B.closed;
    cd.s0 = win ? s0_d : nullptr;
    // (k_refill keeps one int per slot in LDS: windows of up to 8192 slots)
    if (win && op->queue_dev && M.nb <= 8192) {
      const int nch = (int)nl_all.size();
      if (dalloc(&wq_d, 8) || dalloc(&slot_chain_d, M.nbp) || dalloc(&slot_dirty_d, M.nbp) || dalloc(&nl_all_d, nch) ||
          dalloc(&so_all_d, nch) || dalloc(&xslots_d, M.nbp) || dalloc(&xchains_d, M.nbp))
        return 1;
      const int wq0[8] = {qnext, 0, 0, 0, M.nb, 1, 0, 0};
      HCHECK(hipMemcpyAsync(wq_d, wq0, sizeof(wq0), hipMemcpyHostToDevice, stream));
      HCHECK(hipMemcpyAsync(slot_chain_d, slot_chain.data(), M.nbp * sizeof(int), hipMemcpyHostToDevice, stream));
      HCHECK(hipMemsetAsync(slot_dirty_d, 0, M.nbp * sizeof(int), stream));
      HCHECK(hipMemcpyAsync(nl_all_d, nl_all.data(), nch * sizeof(int), hipMemcpyHostToDevice, stream));
      HCHECK(hipMemcpyAsync(so_all_d, so_all.data(), nch * sizeof(int), hipMemcpyHostToDevice, stream));
      HCHECK(hipStreamSynchronize(stream));  // the host sources above
      dev_queue = true;
    }
    sgrid = dim3((unsigned)((nv + 255) / 256));
    if (ed3 && block_start(ocs, sides)) return 1;
    if constexpr (std::is_same<T, cplx>::value) {
      dev_depth = lin && !proj && op->depth_tol > 0.0 && op->depth_tail != 1 && maxdepth <= DEPTH_MAX;
      if (dev_depth) {
        if (dalloc(&gprev_d, M.nbp) || dalloc(&fprev_d, M.nbp) || dalloc(&tnorm_d, M.nbp) || dalloc(&conv_d, M.nbp)) return 1;
        HCHECK(hipMemsetAsync(conv_d, 0, M.nbp * sizeof(int), stream));
        HCHECK(hipStreamCreateWithFlags(&tstream, hipStreamNonBlocking));
        HCHECK(hipEventCreateWithFlags(&ev_a, hipEventDisableTiming));
        HCHECK(hipEventCreateWithFlags(&ev_t, hipEventDisableTiming));
        std::vector<double> tn(M.nbp, 0.0);
        for (size_t q = 0; q < tnorm.size(); ++q) tn[q] = tnorm[q];
        HCHECK(hipMemcpyAsync(tnorm_d, tn.data(), M.nbp * 8, hipMemcpyHostToDevice, stream));
        HCHECK(hipMemsetAsync(gprev_d, 0, M.nbp * sizeof(cplx), stream));
        HCHECK(hipMemsetAsync(fprev_d, 0, M.nbp * sizeof(cplx), stream));
        HCHECK(hipStreamSynchronize(stream));
      }
    }
    return 0;
  }

  // Start vectors of the blocks (ed3_start, src/asr_ed3.f90:131-144; epi_start, src/asr_epi.f90:578-587): vv = h'|s0>
  // from the start vectors in u (one S pass, normalised coordinates), projected; nv^2 = vv.vv on the device,
  // na^2 = s0.s0 + 2 o_c s0.vv + o_c^2 nv^2 (s0 lives on the start node); then u and the projection vector t of each
  // member, and c_1 = t.u.
  int block_start(const std::vector<double>& ocs, const std::vector<int>& sides) {
    if constexpr (!std::is_same<T, cplx>::value) {
      return fail("chain blocks: complex chains only");
    } else {
      const int nb = M.nb, nbp = M.nbp;
      if (!M.uniform_d) return fail("chain blocks: h' needs one d for every species (D_c^1/2)");
      md.tact = tact;
      FArgs f = fargs0();
      f.xs = B.u; f.xs_norm = 1; f.s_din = 1; f.s_dout = 1; f.s_coef = 1.0;
      f.xr[0] = B.u; f.qr[0] = QC; f.r_norm[0] = 1; f.r_coef[0] = 1.0; f.nr = 1;
      f.out = B.w; f.out_sq = 1;
      if (launch_fused<T>(op, md, f, stream)) return 1;
      hipLaunchKernelGGL((k_update<T, false>), grid, dim3(BLOCK), 0, stream, dev_of(op), md, B.w, B.u, B.up, B.a_eff, B.b_eff,
                         project, B.part, (const T*)nullptr, (T*)nullptr, (const T*)nullptr, (const T*)nullptr,
                         (const int*)nullptr);
      hipLaunchKernelGGL(k_sum_member<T>, dim3(nb), dim3(BLOCK), 0, stream, B.part, nbx, md.mt, B.inv_b);
      HCHECK(hipGetLastError());
      std::vector<T> nv2(nbp), row0((size_t)9 * nbp);
      HCHECK(hipMemcpyAsync(nv2.data(), B.inv_b, nbp * sizeof(T), hipMemcpyDeviceToHost, stream));
      HCHECK(hipMemcpyAsync(row0.data(), B.w, row0.size() * sizeof(T), hipMemcpyDeviceToHost, stream));
      HCHECK(hipStreamSynchronize(stream));
      std::vector<double> coef((size_t)4 * nbp, 0.0);
      std::vector<T> c1((size_t)nbp * ld);
      std::memset(c1.data(), 0, c1.size() * sizeof(T));
      snorm.assign(nb, 0.0);
      tnorm.assign(nb, 0.0);
      for (int q = 0; q < nb; ++q) {
        const double oc = ocs[q];
        const double vn2 = nv2[q].re;
        cplx vi{0.0, 0.0};
        double s00 = 0.0;
        for (int l = 0; l < 9; ++l) {
          const double sl = s0h[(size_t)l * nbp + q];
          vi = vi + sl * row0[(size_t)l * nbp + q];
          s00 += sl * sl;
        }
        if (std::fabs(s00 - 1.0) > 1.0e-12) return fail("chain blocks: the start vector must have unit norm");
        const double nvv = std::sqrt(vn2);
        const double na = std::sqrt((s00 + 2.0 * oc * vi.re + oc * oc * vn2));
        if (!(nvv >= 1.0e-14) || !(na >= 1.0e-14)) return fail("ed3_start / epi_start: zero start vector");
        double su, sv, tsu, tsv;
        if (sides[q] == 0) { su = 1.0 / na; sv = oc / na; tsu = 0.0; tsv = 1.0; snorm[q] = na; tnorm[q] = nvv; }
        else { su = 0.0; sv = 1.0 / nvv; tsu = 1.0; tsv = oc; snorm[q] = nvv; tnorm[q] = na; }
        coef[4 * q] = su; coef[4 * q + 1] = sv; coef[4 * q + 2] = tsu; coef[4 * q + 3] = tsv;
        // c_1 = t.u = tsu su s0.s0 + (tsu sv + tsv su) s0.vv + tsv sv vv.vv
        const cplx vv2 = nv2[q];
        c1[(size_t)q * ld] = cplx{tsu * su * s00, 0.0} + (tsu * sv + tsv * su) * vi + (tsv * sv) * vv2;
      }
      double *dcoef = nullptr, *ds0 = nullptr;
      if (dalloc(&dcoef, coef.size()) || dalloc(&ds0, s0h.size())) return 1;
      int st = 0;
      st |= hipMemcpyAsync(dcoef, coef.data(), coef.size() * 8, hipMemcpyHostToDevice, stream) != hipSuccess;
      st |= hipMemcpyAsync(ds0, s0h.data(), s0h.size() * 8, hipMemcpyHostToDevice, stream) != hipSuccess;
      st |= hipMemcpyAsync(B.c, c1.data(), c1.size() * sizeof(T), hipMemcpyHostToDevice, stream) != hipSuccess;
      if (!st) {
        hipLaunchKernelGGL(k_block_start, sgrid, dim3(256), 0, stream, (const cplx*)B.w, (cplx*)B.u, (cplx*)B.tv,
                           (const double*)dcoef, (const double*)ds0, nv, nbp);
        st |= hipGetLastError() != hipSuccess;
      }
      st |= hipMemsetAsync(B.w, 0, nv * sizeof(T), stream) != hipSuccess;
      st |= hipMemsetAsync(B.part, 0, (size_t)nbx * M.nbp * sizeof(T), stream) != hipSuccess;
      st |= hipMemsetAsync(B.inv_b, 0, M.nbp * sizeof(T), stream) != hipSuccess;
      st |= hipMemsetAsync(B.closed, 0, M.nbp * sizeof(int), stream) != hipSuccess;
      st |= hipStreamSynchronize(stream) != hipSuccess;
      (void)hipFree(dcoef);
      (void)hipFree(ds0);
      if (st) return fail("chain blocks: start vectors");
      return 0;
    }
  }

  // the launches of one level on the group's stream (no host synchronisation): the operator pass and a_k, then the
  // update, b_k and the scaling.  With the adaptive depth the host looks at a_k between the two (adaptive_check).
  int step() {
    if (step_apply()) return 1;
    if (dev_depth) {
      if (depth_test_dev()) return 1;
    } else if (adaptive_check()) {
      return 1;
    }
    return step_update();
  }

  // the adaptive depth test of this level on the device (k_depth_test): at levels 8, 12, 16, ... (1-based)
  int depth_test_dev() {
    if constexpr (!std::is_same<T, cplx>::value) {
      return 0;
    } else {
      const int k1 = level_host + 1;
      if (!win && (k1 < 8 || (k1 % 4) != 0)) return 0;
      if (win && !dev_queue) {
        // window runs: the kernel takes each member's own level count; no launch when no member is at a test level (with
        // the queue on the device the host does not follow the slots, and the kernel runs every level)
        bool any = false;
        for (int q = 0; q < M.nb && !any; ++q) {
          const int c = slot_chain[q];
          if (c < 0) continue;
          const int rel = k1 - slot_s0[q];
          any = rel >= 8 && rel % 4 == 0 && rel < nl_all[c];
        }
        if (!any) return 0;
      }
      // on the test stream, after this level's a_k (k_fin_a) and beside its update; joined before the next k_fin_a
      HCHECK(hipEventRecord(ev_a, stream));
      HCHECK(hipStreamWaitEvent(tstream, ev_a, 0));
      hipLaunchKernelGGL(k_depth_test, dim3((M.nb + DT_TPB - 1) / DT_TPB), dim3(DT_TPB), 0, tstream, cd, M.nb,
                         (const cplx*)B.a, (const cplx*)B.b, ed3 ? (const cplx*)B.c : (const cplx*)nullptr,
                         (const cplx*)md.z, k1, op->depth_tol, op->depth_tail, gprev_d, fprev_d, (const double*)tnorm_d,
                         conv_d);
      HCHECK(hipGetLastError());
      HCHECK(hipEventRecord(ev_t, tstream));
      test_pending = true;
      return 0;
    }
  }

  // the main stream waits for the last depth test (before a k_fin_a, a compaction, or the end of the run)
  int join_test() {
    if (!test_pending) return 0;
    HCHECK(hipStreamWaitEvent(stream, ev_t, 0));
    test_pending = false;
    return 0;
  }
  // the end of a run: the decisions of the last test take effect
  int flush_test() {
    if (!test_pending) return 0;
    if (join_test()) return 1;
    hipLaunchKernelGGL(k_conv_apply<T>, dim3((M.nb + 63) / 64), dim3(64), 0, stream, cd, M.nb, conv_d, B.b);
    HCHECK(hipGetLastError());
    return 0;
  }

  int step_apply() {
    md.tact = tact;
    if (g_timer) g_timer->ord = 0;
    if (reorth) HCHECK(hipMemcpyAsync(B.qv + (size_t)level_host * nv, B.u, nv * sizeof(double), hipMemcpyDeviceToDevice, stream));
    if (proj) {
      const int nt = op->ntarget;
      hipLaunchKernelGGL(k_proj<T>, dim3((M.nb + 255) / 256, nt * 9), dim3(256), 0, stream, dev_of(op), md, (const T*)B.u, nt,
                         (const int*)op->tnode, (const int*)op->tgop, (const double*)B.tw,
                         B.cpr + (size_t)level_host * nt * 9 * M.nbp, s_cur());
      HCHECK(hipGetLastError());
    }
    if (fused) {
      if (fused_apply()) return 1;
    } else if (apply_level_op<T>(op, M, md, B.u, B.w, B.t1, B.t2, B.h1, lin, B.u, B.part, stream)) {
      return 1;
    }
    if (join_test()) return 1;
    if (timer_mark(stream, KIND_FIN_A, true)) return 1;
    hipLaunchKernelGGL(k_fin_a<T>, dim3(M.nb), dim3(BLOCK), 0, stream, cd, M.nb, B.part, nbx, apply_mt, B.a, B.a_eff, s_cur(),
                       (const T*)B.lpart, fused ? op->nplist : 0, M.nbp, conv_d, B.b);
    if (timer_mark(stream, 0, false)) return 1;
    return 0;
  }

  // the fused level, first half: h1 = rand_c(u~) + D S D u~, t1 = X_o h1 and the partials of a_k (tile pass at the nodes
  // without flip partners, k_list_terms at the others).  Route ed3: h1 = h' u~ and t1 = t = u u~ - Qt h1 (the first pass
  // of apply_ed3_op) with the partials of its a_k.
  int fused_apply() {
    if constexpr (!std::is_same<T, cplx>::value) {
      return fail("fused level: complex chains only");
    } else {
      FArgs f = fargs0();
      f.xs = B.u; f.xs_norm = 1; f.s_din = 1; f.s_dout = 1; f.s_coef = 1.0;
      f.xr[0] = B.u; f.qr[0] = QC; f.r_norm[0] = 1; f.r_coef[0] = 1.0; f.nr = 1;
      f.out = B.h1;
      const bool has_t1 = M.ed3 || !M.zero_o;
      if (M.ed3) {
        f.out2 = B.t1; f.q2 = QQA; f.c2 = -1.0;
        f.x2 = B.u; f.q2x = QUA; f.x2norm = 1; f.c2x = 1.0;
        f.fold = 1;
      } else if (!M.zero_o) {
        f.out2 = B.t1; f.q2 = QO; f.c2 = 1.0;
        f.fold = 1;
      }
      f.apart = 1; f.a_e = (M.ed3 || M.zero_e) ? 0 : 1; f.uo = B.u; f.partial = B.part;
      if (launch_fused<T>(op, md, f, stream)) return 1;
      if (op->nplist > 0) {
        if (timer_mark(stream, KIND_FIX, true)) return 1;
        const size_t nt = (size_t)op->nplist * md.nbp;
        hipLaunchKernelGGL(k_list_terms, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, stream, dev_of(op), md,
                           (const int*)op->plist, op->nplist, (const cplx*)B.h1, has_t1 ? (cplx*)B.t1 : (cplx*)nullptr,
                           (const cplx*)B.u, f.a_e, (cplx*)B.lpart, M.ed3 ? 1 : 0);
        HCHECK(hipGetLastError());
        if (timer_mark(stream, 0, false)) return 1;
      }
      return 0;
    }
  }

  // the fused level, second half: w = sq (E u~ + h1 - h t1 + X(z) u~) - a u - b u_prev, projected, and <w, w>; route ed3
  // with sq [(z - et) u~ + h1 - u h1 - h'(t)] (the last pass of apply_ed3_op); the chain blocks also <tv, w>
  int fused_update() {
    FArgs f = fargs0();
    if (M.ed3) {
      f.xs = B.t1; f.xs_norm = 0; f.s_din = 1; f.s_dout = 1; f.s_coef = -1.0;
      f.xr[0] = B.t1; f.qr[0] = QC; f.r_norm[0] = 0; f.r_coef[0] = -1.0; f.nr = 1;
      f.xl[0] = B.u; f.lq[0] = QPA; f.l_norm[0] = 1; f.l_coef[0] = 1.0;
      f.xl[1] = B.h1; f.lq[1] = QUA; f.l_norm[1] = 0; f.l_coef[1] = -1.0; f.nl = 2;
    } else {
      int nr = 0;
      if (!M.zero_o) {
        f.xs = B.t1; f.s_din = 1; f.s_dout = 1; f.s_coef = -1.0;
        f.xr[nr] = B.t1; f.qr[nr] = QC; f.r_norm[nr] = 0; f.r_coef[nr] = -1.0; ++nr;
      }
      if (!M.zero_e) { f.xr[nr] = B.u; f.qr[nr] = QE; f.r_norm[nr] = 1; f.r_coef[nr] = 1.0; ++nr; }
      f.nr = nr;
      f.xl[0] = B.u; f.lq[0] = QPA; f.l_norm[0] = 1; f.l_coef[0] = 1.0; f.nl = 1;
    }
    f.xo[0] = B.h1; f.o_coef[0] = 1.0; f.no = 1;
    f.out_sq = 1;
    f.upd = 1; f.proj = project; f.uo = B.u; f.upv = B.up; f.aeff = B.a_eff; f.beff = B.b_eff; f.scv = s_cur();
    f.scpv = s_prev(); f.act = B.active;
    f.out = B.w; f.partial = B.part; f.partial2 = B.pchk;
    f.fold = (M.ed3 || !M.zero_o) ? 2 : 0;  // h1 holds g at the nodes without flip partners (the first pass folded)
    if (ed3) { f.tvv = B.tv; f.partial3 = B.part2; }
    return launch_fused<T>(op, md, f, stream);
  }

  // Adaptive depth of edc_green (src/asr_edc.f90:121-127): at levels k = 8, 12, 16, ... (1-based) of a chain that has
  // not reached its depth, G(z) of the chain terminated at k (cf_species with the operator's tail); from k = 12 on the
  // chain leaves when |G_k - G_{k-4}| <= depth_tol |G_k|, keeping a_1..a_k and b_1..b_{k-1}.  Route ed3 (chain,
  // src/asr_ed3.f90:177-184) also asks |f_k - f_{k-4}| <= depth_tol max(|f_k|, |t| |G_k|) of the projection f.  One host
  // synchronisation per tested level; the coefficients are a few KB.
  int adaptive_check() {
    if constexpr (!std::is_same<T, cplx>::value) return 0;
    else {
      const int k1 = level_host + 1;
      // the projected chains run all levels (edc_green_proj has no depth test)
      if (proj || !(op->depth_tol > 0.0) || !lin || k1 < 8 || (k1 % 4) != 0) return 0;
      const int nb = M.nb;
      std::vector<cplx> ah((size_t)nb * ld), bh((size_t)nb * ld), ch(ed3 ? (size_t)nb * ld : 0);
      std::vector<int> act(nb), dep(nb);
      HCHECK(hipMemcpyAsync(ah.data(), B.a, ah.size() * sizeof(cplx), hipMemcpyDeviceToHost, stream));
      HCHECK(hipMemcpyAsync(bh.data(), B.b, bh.size() * sizeof(cplx), hipMemcpyDeviceToHost, stream));
      if (ed3) HCHECK(hipMemcpyAsync(ch.data(), B.c, ch.size() * sizeof(cplx), hipMemcpyDeviceToHost, stream));
      HCHECK(hipMemcpyAsync(act.data(), B.active, nb * sizeof(int), hipMemcpyDeviceToHost, stream));
      HCHECK(hipMemcpyAsync(dep.data(), B.depth, nb * sizeof(int), hipMemcpyDeviceToHost, stream));
      HCHECK(hipStreamSynchronize(stream));
      if (gprev.size() != (size_t)nb) gprev.assign(nb, cplx{0.0, 0.0});
      if (fprev.size() != (size_t)nb) fprev.assign(nb, cplx{0.0, 0.0});
      std::vector<double> cz(2 * k1, 0.0);
      // G_k and f_k of every member still running, on a few host threads (independent continued fractions)
      std::vector<int> idx;
      for (int q = 0; q < nb; ++q)
        if (act[q] && k1 < dep[q]) idx.push_back(q);
      std::vector<cplx> gv(nb), fv(nb);
      auto eval = [&](size_t lo, size_t hi) {
        for (size_t t = lo; t < hi; ++t) {
          const int q = idx[t];
          double g[2], f[2];
          const double zz[2] = {M.z[q].re, M.z[q].im};
          const double* cq = ed3 ? (const double*)&ch[(size_t)q * ld] : cz.data();
          // the status counts bp fallbacks (hull values then), not errors
          (void)asr_host_cf_species((const double*)&ah[(size_t)q * ld], (const double*)&bh[(size_t)q * ld], cq, k1, zz,
                                    op->depth_tail, g, f, nullptr, nullptr);
          gv[q] = cplx{g[0], g[1]};
          fv[q] = cplx{f[0], f[1]};
        }
      };
      const size_t nidx = idx.size();
      const int nth = (int)std::min<size_t>(std::max(1u, std::min(16u, std::thread::hardware_concurrency())), nidx / 16 + 1);
      if (nth > 1) {
        std::vector<std::thread> th;
        for (int w = 0; w < nth; ++w) th.emplace_back(eval, nidx * w / nth, nidx * (w + 1) / nth);
        for (auto& t : th) t.join();
      } else {
        eval(0, nidx);
      }
      bool changed = false;
      for (const int q : idx) {
        const cplx gt = gv[q], ft = fv[q];
        if (k1 > 8) {
          const cplx dg = gt - gprev[q], df = ft - fprev[q];
          const double ag = std::hypot(gt.re, gt.im);
          bool conv = std::hypot(dg.re, dg.im) <= op->depth_tol * ag;
          if (ed3)
            conv = conv && std::hypot(df.re, df.im) <= op->depth_tol * std::max(std::hypot(ft.re, ft.im), tnorm[q] * ag);
          if (conv) {
            act[q] = 0;
            changed = true;
          }
        }
        gprev[q] = gt;
        fprev[q] = ft;
      }
      if (changed) {
        HCHECK(hipMemcpyAsync(B.active, act.data(), nb * sizeof(int), hipMemcpyHostToDevice, stream));
        HCHECK(hipStreamSynchronize(stream));
      }
      return 0;
    }
  }

  int step_update() {
    if (fused) {
      if (fused_update()) return 1;
    } else {
    if (timer_mark(stream, KIND_UPDATE, true)) return 1;
    if (md.mt * (int)sizeof(T) >= 128)
      hipLaunchKernelGGL((k_update<T, true>), grid, dim3(BLOCK), 0, stream, dev_of(op), md, B.w, B.u, B.up, B.a_eff, B.b_eff,
                         project, B.part, (const T*)B.tv, B.part2, s_cur(), s_prev(),
                         lag ? (const int*)B.active : (const int*)nullptr);
    else
      hipLaunchKernelGGL((k_update<T, false>), grid, dim3(BLOCK), 0, stream, dev_of(op), md, B.w, B.u, B.up, B.a_eff, B.b_eff,
                         project, B.part, (const T*)B.tv, B.part2, s_cur(), s_prev(),
                         lag ? (const int*)B.active : (const int*)nullptr);
    if (reorth) {
      // w = w - Q (Q^T w), twice (classical Gram-Schmidt against q_0..q_k); the second pass leaves the partials of w w
      const int nq = level_host + 1;
      for (int pass = 0; pass < 2; ++pass) {
        hipLaunchKernelGGL(k_qdots, grid, dim3(BLOCK), 0, stream, dev_of(op), md, (const double*)B.w, (const double*)B.qv, nv,
                           nq, B.qpart, qplane);
        hipLaunchKernelGGL(k_qfin, dim3(M.nb), dim3(BLOCK), 0, stream, (const double*)B.qpart, qplane, nbx, md.mt, nq, M.nbp,
                           B.qcoef);
        hipLaunchKernelGGL(k_qsub, grid, dim3(BLOCK), 0, stream, dev_of(op), md, (double*)B.w, (const double*)B.qv, nv, nq,
                           (const double*)B.qcoef, pass == 1 ? (double*)B.part : (double*)nullptr);
      }
    }
    if (timer_mark(stream, 0, false)) return 1;
    }
    if (timer_mark(stream, KIND_FIN_B, true)) return 1;
    hipLaunchKernelGGL(k_fin_b<T>, dim3(M.nb), dim3(BLOCK), 0, stream, cd, M.nb, B.part, nbx, md.mt, B.b, B.b_eff,
                       B.inv_b, (const T*)B.part2, s_cur(), s_next(), (const T*)B.pchk, B.asym, (const T*)B.a);
    if (timer_mark(stream, 0, false)) return 1;
    if (!lag) {
      if (timer_mark(stream, KIND_SCALE, true)) return 1;
      hipLaunchKernelGGL(k_scale<T>, sgrid, dim3(256), 0, stream, B.w, B.inv_b, nv, M.nbp);
      if (timer_mark(stream, 0, false)) return 1;
    } else if (level_host % 16 == 15) {
      // normalise once every 16 levels, so that the scale (a product of b's) stays far from under- and overflow
      if (timer_mark(stream, KIND_SCALE, true)) return 1;
      hipLaunchKernelGGL(k_rescale<T>, sgrid, dim3(256), 0, stream, B.w, (const T*)s_next(), nv, M.nbp);
      hipLaunchKernelGGL(k_scale_one<T>, dim3((M.nbp + 255) / 256), dim3(256), 0, stream, s_next(), M.nbp);
      if (timer_mark(stream, 0, false)) return 1;
    }
    hipLaunchKernelGGL(k_level_end, dim3(1), dim3(64), 0, stream, B.level, B.active, M.nb, M.mt, tact, ntile);
    HCHECK(hipGetLastError());
    if (g_timer) {
      HCHECK(hipStreamSynchronize(stream));
      if (g_acc.size() < (size_t)g_timer->ord) {
        g_acc.resize(g_timer->ord, 0.0);
        g_cnt.resize(g_timer->ord, 0);
      }
      for (int q = 0; q < g_timer->ord; ++q) {
        float t = 0.f;
        HCHECK(hipEventElapsedTime(&t, g_timer->ev[2 * q], g_timer->ev[2 * q + 1]));
        g_acc[q] += t;
        g_cnt[q] += 1;
      }
      g_levels++;
    }
    std::swap(B.up, B.u);  // up <- u
    std::swap(B.u, B.w);   // u <- w / b ; w <- old up (overwritten next level)
    ++level_host;
    return 0;
  }

  // Member compaction between host-issued levels (the reference's edc_green_blk lets every member leave at its own depth;
  // here a member tile runs while any of its members does).  After a level of the adaptive-depth test (levels 12, 16, ...)
  // the members still running move to the front, in order, and those that have left behind them, when that frees at
  // least one member tile for at least two more levels.  Every per-member array moves alike: the chain vectors u, u_prev
  // (and route ed3's projection vector), the member tables of the operator pass, the coefficients a, b (c) with the
  // scalars of the recursion, the lagged scales and the depth-test state, and on the host the chain order and the
  // per-member values the host reads.  *alive = 0 when no member runs any more.
  bool can_compact() const { return lag && lin && !proj && !reorth && op->depth_tol > 0.0; }

  template <class E>
  int perm_cols(E* arr, size_t nrows) {
    const size_t bytes = nrows * (size_t)M.nbp * sizeof(E);
    if (bytes > ptmp_bytes) {
      if (ptmp) (void)hipFree(ptmp);
      ptmp = nullptr;
      ptmp_bytes = 0;
      HCHECK(hipMalloc(&ptmp, bytes));
      ptmp_bytes = bytes;
    }
    HCHECK(hipMemcpyAsync(ptmp, arr, bytes, hipMemcpyDeviceToDevice, stream));
    const size_t n = nrows * (size_t)M.nbp;
    const unsigned nblk = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_perm_cols<E>, dim3(nblk), dim3(256), 0, stream, arr, (const E*)ptmp, nrows, M.nbp, M.nbp,
                       (const int*)perm_d);
    HCHECK(hipGetLastError());
    return 0;
  }
  template <class E>
  int perm_rows(E* arr) {
    const size_t bytes = (size_t)M.nbp * ld * sizeof(E);
    if (bytes > ptmp_bytes) {
      if (ptmp) (void)hipFree(ptmp);
      ptmp = nullptr;
      ptmp_bytes = 0;
      HCHECK(hipMalloc(&ptmp, bytes));
      ptmp_bytes = bytes;
    }
    HCHECK(hipMemcpyAsync(ptmp, arr, bytes, hipMemcpyDeviceToDevice, stream));
    const size_t n = (size_t)M.nbp * ld;
    hipLaunchKernelGGL(k_perm_rows<E>, dim3((unsigned)std::min<size_t>((n + 255) / 256, 8192)), dim3(256), 0, stream, arr,
                       (const E*)ptmp, ld, M.nbp, (const int*)perm_d);
    HCHECK(hipGetLastError());
    return 0;
  }
  // a chain vector into a free level buffer, which then takes its place: only the slots of the member tiles that still
  // run (ncol) are written, the vectors of members that have left are not needed any more
  int perm_vector(T*& v, T*& spare, int ncol) {
    const size_t n = nv / M.nbp * ncol;
    hipLaunchKernelGGL(k_perm_cols<T>, dim3((unsigned)std::min<size_t>((n + 255) / 256, 16384)), dim3(256), 0, stream, spare,
                       (const T*)v, nv / M.nbp, M.nbp, ncol, (const int*)perm_d);
    HCHECK(hipGetLastError());
    std::swap(v, spare);
    return 0;
  }

  int compact(int* alive) {
    *alive = 1;
    if (win) return refill(alive);
    // members leave at the test levels 12, 16, ...; with the test on the device its decision takes effect one level later
    const int ph = dev_depth ? 1 : 0;
    if (!can_compact() || level_host < 12 + ph || level_host % 4 != ph) return 0;
    if (test_pending) HCHECK(hipEventSynchronize(ev_t));
    const int nb = M.nb, mt = M.mt;
    std::vector<int> act(nb);
    HCHECK(hipMemcpyAsync(act.data(), B.active, nb * sizeof(int), hipMemcpyDeviceToHost, stream));
    HCHECK(hipStreamSynchronize(stream));
    int nlive = 0;
    for (int q = 0; q < nb; ++q) nlive += act[q] ? 1 : 0;
    if (nlive == 0) {
      *alive = 0;
      return 0;
    }
    return compact_live(act, nlive, maxdepth - level_host);
  }

  // the live members (act) to the front when that frees member-tile levels (remaining: levels still to run)
  int compact_live(const std::vector<int>& act, int nlive, int remaining) {
    const int nb = M.nb, mt = M.mt;
    int cur = 0;
    for (int t = 0; t * mt < nb; ++t) {
      int any = 0;
      for (int q = t * mt; q < std::min(nb, (t + 1) * mt); ++q) any |= act[q];
      cur += any ? 1 : 0;
    }
    const int need = (nlive + mt - 1) / mt;
    // worth a vector gather when it frees member-tile levels: at least one tile for four more levels or two for two
    if (need >= cur || (cur - need) * remaining < 4) return 0;
    // stable partition: the running members first, then those that have left; padding slots stay
    std::vector<int> map(M.nbp);
    int w = 0;
    for (int q = 0; q < nb; ++q)
      if (act[q]) map[w++] = q;
    for (int q = 0; q < nb; ++q)
      if (!act[q]) map[w++] = q;
    for (int q = nb; q < M.nbp; ++q) map[q] = q;
    if (!perm_d && dalloc(&perm_d, M.nbp)) return 1;
    HCHECK(hipMemcpyAsync(perm_d, map.data(), M.nbp * sizeof(int), hipMemcpyHostToDevice, stream));
    // chain vectors (w, t1, t2 are free between levels)
    if (perm_vector(B.u, B.w, need * mt) || perm_vector(B.up, B.t2, need * mt)) return 1;
    if (B.tv && perm_vector(B.tv, B.t1, need * mt)) return 1;
    // member tables of the operator pass
    const size_t nbp = M.nbp;
    if (perm_cols(MB.chim, 1) || perm_cols(MB.wts, M.wts.size() / nbp) || perm_cols(MB.fpar, M.fpar.size() / nbp) ||
        perm_cols(MB.dvec, M.dvec.size() / nbp) || perm_cols(MB.dperm, M.dperm.size() / nbp) ||
        perm_cols(MB.mtm, M.mtm.size() / nbp) || perm_cols(MB.fsp, M.fsp.size() / nbp) || perm_cols(MB.z, 1))
      return 1;
    // the recursion's per-member state
    if (perm_rows(B.a) || perm_rows(B.b) || (B.c && perm_rows(B.c))) return 1;
    for (T* p : {B.a_eff, B.b_eff, B.inv_b, sc3[0], sc3[1], sc3[2]})
      if (p && perm_cols(p, 1)) return 1;
    for (int* p : {B.depth, B.active, B.mused, B.closed, s0_d, conv_d})
      if (p && perm_cols(p, 1)) return 1;
    if constexpr (std::is_same<T, cplx>::value) {
      if (gprev_d && (perm_cols(gprev_d, 1) || perm_cols(fprev_d, 1) || perm_cols(tnorm_d, 1))) return 1;
    }
    // the tiles with running members are the first `need`: later levels launch those alone (member tiles below a wave;
    // the lane-pair form of tile 64 keeps its full grid)
    std::vector<int> ta(std::max(ntile, 2), 0);
    for (int t = 0; t < need; ++t) ta[t] = 1;
    HCHECK(hipMemcpyAsync(tact, ta.data(), ta.size() * sizeof(int), hipMemcpyHostToDevice, stream));
    if (M.mt < WAVE) {
      md.ntl = need;
      grid = node_grid(op, md);
    }
    // host bookkeeping (a window run keeps the chain order of its results; its slots move)
    auto permute = [&](auto& v) {
      if (v.size() < (size_t)nb) return;
      auto old = v;
      for (int q = 0; q < nb; ++q) v[q] = old[map[q]];
    };
    if (!win) permute(order);
    permute(M.z);
    permute(gprev);
    permute(fprev);
    permute(snorm);
    permute(tnorm);
    permute(slot_chain);
    permute(slot_s0);
    permute(slot_dirty);
    HCHECK(hipStreamSynchronize(stream));  // the host sources above and the staging buffer
    ++ncompact;
    ++op->ncompact;
    return 0;
  }

  // Window runs, after every level: the members that may have left by now (after their own test levels 12, 16, ... of
  // the adaptive depth, or at their depth) are looked at; those that left hand their results to their chains' rows and
  // their slots to the next chains of the queue; once the queue is empty the live members are compacted as above.
  int refill(int* alive) {
    const int nb = M.nb, mt = M.mt;
    bool maybe = false;
    for (int q = 0; q < nb && !maybe; ++q) {
      const int c = slot_chain[q];
      if (c < 0) continue;
      const int rel = level_host - slot_s0[q];
      // at its depth, or after its test levels 12, 16, ... (a member the test finds converged leaves at once)
      if (rel >= nl_all[c] || (rel >= 12 && rel % 4 == 0)) maybe = true;
    }
    if (!maybe) return 0;
    if (join_test()) return 1;
    hipLaunchKernelGGL(k_conv_leave<T>, dim3((nb + 63) / 64), dim3(64), 0, stream, cd, nb, conv_d, B.b);
    HCHECK(hipGetLastError());
    std::vector<int> act(nb), cls(nb);
    HCHECK(hipMemcpyAsync(act.data(), B.active, nb * sizeof(int), hipMemcpyDeviceToHost, stream));
    HCHECK(hipMemcpyAsync(cls.data(), B.closed, nb * sizeof(int), hipMemcpyDeviceToHost, stream));
    HCHECK(hipStreamSynchronize(stream));
    std::vector<int> xs, xc;
    for (int q = 0; q < nb; ++q)
      if (slot_chain[q] >= 0 && !act[q]) {
        xs.push_back(q);
        xc.push_back(slot_chain[q]);
        slot_chain[q] = -1;
        if (cls[q]) slot_dirty[q] = 1;
      }
    if (!xs.empty()) {
      const int n = (int)xs.size();
      HCHECK(hipMemcpyAsync(wslots, xs.data(), n * sizeof(int), hipMemcpyHostToDevice, stream));
      HCHECK(hipMemcpyAsync(wchains, xc.data(), n * sizeof(int), hipMemcpyHostToDevice, stream));
      hipLaunchKernelGGL(k_extract<T>, dim3(n), dim3(64), 0, stream, a_all, b_all, mused_all, (const T*)B.a, (const T*)B.b,
                         (const int*)B.mused, ld, (const int*)wslots, (const int*)wchains, n);
      HCHECK(hipGetLastError());
      HCHECK(hipStreamSynchronize(stream));  // before the slot lists are reused
    }
    std::vector<int> ns, nc, nd, nl0;
    int ndirty = 0;
    for (int q = 0; q < nb && qnext < nchain; ++q)
      if (slot_chain[q] < 0) {
        const int c = qnext++;
        slot_chain[q] = c;
        slot_s0[q] = level_host;
        act[q] = nl_all[c] > 0 ? 1 : 0;
        ns.push_back(q);
        nc.push_back(c);
        nd.push_back(nl_all[c]);
        nl0.push_back(so_all[c]);
        ndirty += slot_dirty[q];
      }
    // slots with a non-zero vector (the chain before ended early) first, for the full start vector; then the clean ones
    std::stable_partition(ns.begin(), ns.end(), [&](int q) { return slot_dirty[q] != 0; });
    {
      std::vector<int> c2, d2, l2;
      for (int q : ns) {
        c2.push_back(slot_chain[q]);
        d2.push_back(nl_all[slot_chain[q]]);
        l2.push_back(so_all[slot_chain[q]]);
      }
      nc.swap(c2);
      nd.swap(d2);
      nl0.swap(l2);
      for (int q : ns) slot_dirty[q] = 0;
    }
    if (!ns.empty()) {
      const int n = (int)ns.size();
      HCHECK(hipMemcpyAsync(wslots, ns.data(), n * sizeof(int), hipMemcpyHostToDevice, stream));
      HCHECK(hipMemcpyAsync(wchains, nc.data(), n * sizeof(int), hipMemcpyHostToDevice, stream));
      HCHECK(hipMemcpyAsync(wvals, nd.data(), n * sizeof(int), hipMemcpyHostToDevice, stream));
      HCHECK(hipMemcpyAsync(wl0, nl0.data(), n * sizeof(int), hipMemcpyHostToDevice, stream));
      const size_t nbp = M.nbp;
      auto fill = [&](auto* dst, const auto* src, size_t nrows) -> int {
        using E = std::remove_const_t<std::remove_pointer_t<decltype(src)>>;
        const size_t tot = nrows * n;
        hipLaunchKernelGGL(k_fill_cols<E>, dim3((unsigned)std::min<size_t>((tot + 255) / 256, 4096)), dim3(256), 0, stream,
                           (E*)dst, (const E*)src, nrows, M.nbp, Mall.nbp, (const int*)wslots, (const int*)wchains, n);
        HCHECK(hipGetLastError());
        return 0;
      };
      if (fill(MB.chim, MBall.chim, 1) || fill(MB.wts, MBall.wts, M.wts.size() / nbp) ||
          fill(MB.fpar, MBall.fpar, M.fpar.size() / nbp) || fill(MB.dvec, MBall.dvec, M.dvec.size() / nbp) ||
          fill(MB.dperm, MBall.dperm, M.dperm.size() / nbp) || fill(MB.mtm, MBall.mtm, M.mtm.size() / nbp) ||
          fill(MB.fsp, MBall.fsp, M.fsp.size() / nbp) || fill(MB.z, MBall.z, 1))
        return 1;
      const size_t rows = nv / nbp;
      if (ndirty > 0) {
        const size_t tot = rows * ndirty;
        hipLaunchKernelGGL(k_start_cols<T>, dim3((unsigned)std::min<size_t>((tot + 255) / 256, 16384)), dim3(256), 0, stream,
                           B.u, rows, M.nbp, (const int*)wslots, (const int*)wl0, ndirty);
      }
      if (n > ndirty)
        hipLaunchKernelGGL(k_start_unit<T>, dim3((unsigned)((n - ndirty + 63) / 64)), dim3(64), 0, stream, B.u, M.nbp,
                           (const int*)wslots + ndirty, (const int*)wl0 + ndirty, n - ndirty);
      cplx *gp = nullptr, *fp = nullptr;
      if constexpr (std::is_same<T, cplx>::value) {
        gp = gprev_d;
        fp = fprev_d;
      }
      hipLaunchKernelGGL(k_member_start<T>, dim3(n), dim3(64), 0, stream, cd, (const int*)wslots, (const int*)wvals, n,
                         level_host, B.a, B.b, B.a_eff, B.b_eff, B.inv_b, sc3[0], sc3[1], sc3[2], gp, fp, s0_d,
                         B.closed, conv_d);
      HCHECK(hipGetLastError());
    }
    int nlive = 0, remaining = 0;
    for (int q = 0; q < nb; ++q)
      if (slot_chain[q] >= 0 && act[q]) {
        ++nlive;
        remaining = std::max(remaining, nl_all[slot_chain[q]] - (level_host - slot_s0[q]));
      }
    if (nlive == 0 && qnext >= nchain) {
      HCHECK(hipStreamSynchronize(stream));
      *alive = 0;
      return 0;
    }
    std::vector<int> live(nb, 0);
    for (int q = 0; q < nb; ++q) live[q] = (slot_chain[q] >= 0 && act[q]) ? 1 : 0;
    if (qnext >= nchain && op->compact) {
      // the queue is empty: compaction as for a batch
      const int before = ncompact;
      if (compact_live(live, nlive, remaining)) return 1;
      if (ncompact != before) return 0;
    }
    // the member tiles with running members
    std::vector<int> ta(std::max(ntile, 2), 0);
    for (int q = 0; q < nb; ++q)
      if (live[q]) ta[q / mt] = 1;
    HCHECK(hipMemcpyAsync(tact, ta.data(), ta.size() * sizeof(int), hipMemcpyHostToDevice, stream));
    HCHECK(hipStreamSynchronize(stream));  // the host sources above
    return 0;
  }

  // the queue on the device after a level: k_refill, then the same steps as refill with the counts read on the device
  int refill_dev() {
    if (join_test()) return 1;  // this level's depth test writes conv, which k_member_start clears for a new member
    const int nb = M.nb, mt = M.mt, nch = (int)nl_all.size();
    hipLaunchKernelGGL(k_conv_leave<T>, dim3((nb + 63) / 64), dim3(64), 0, stream, cd, nb, conv_d, B.b);
    hipLaunchKernelGGL(k_refill, dim3(1), dim3(RF_TPB), (size_t)nb * sizeof(int), stream, cd, nb, mt, nch, wq_d, slot_chain_d,
                       slot_dirty_d,
                       (const int*)nl_all_d, (const int*)so_all_d, xslots_d, xchains_d, wslots, wchains, wvals, wl0, tact,
                       ntile);
    hipLaunchKernelGGL(k_extract<T>, dim3(nb), dim3(64), 0, stream, a_all, b_all, mused_all, (const T*)B.a, (const T*)B.b,
                       (const int*)B.mused, ld, (const int*)xslots_d, (const int*)xchains_d, 0, (const int*)wq_d + 1);
    const size_t nbp = M.nbp;
    auto fill = [&](auto* dst, const auto* src, size_t nrows) -> int {
      using E = std::remove_const_t<std::remove_pointer_t<decltype(src)>>;
      hipLaunchKernelGGL(k_fill_cols<E>, dim3((unsigned)std::min<size_t>((nrows * nb + 255) / 256, 4096)), dim3(256), 0,
                         stream, (E*)dst, (const E*)src, nrows, M.nbp, Mall.nbp, (const int*)wslots, (const int*)wchains, 0,
                         (const int*)wq_d + 3);
      HCHECK(hipGetLastError());
      return 0;
    };
    if (fill(MB.chim, MBall.chim, 1) || fill(MB.wts, MBall.wts, M.wts.size() / nbp) ||
        fill(MB.fpar, MBall.fpar, M.fpar.size() / nbp) || fill(MB.dvec, MBall.dvec, M.dvec.size() / nbp) ||
        fill(MB.dperm, MBall.dperm, M.dperm.size() / nbp) || fill(MB.mtm, MBall.mtm, M.mtm.size() / nbp) ||
        fill(MB.fsp, MBall.fsp, M.fsp.size() / nbp) || fill(MB.z, MBall.z, 1))
      return 1;
    const size_t rows = nv / nbp;
    hipLaunchKernelGGL(k_start_cols<T>, dim3((unsigned)std::min<size_t>((rows * nb + 255) / 256, 16384)), dim3(256), 0, stream,
                       B.u, rows, M.nbp, (const int*)wslots, (const int*)wl0, 0, (const int*)wq_d + 2);
    hipLaunchKernelGGL(k_start_unit<T>, dim3((unsigned)((nb + 63) / 64)), dim3(64), 0, stream, B.u, M.nbp,
                       (const int*)wslots, (const int*)wl0, 0, (const int*)wq_d + 2);
    cplx *gp = nullptr, *fp = nullptr;
    if constexpr (std::is_same<T, cplx>::value) {
      gp = gprev_d;
      fp = fprev_d;
    }
    hipLaunchKernelGGL(k_member_start<T>, dim3(nb), dim3(64), 0, stream, cd, (const int*)wslots, (const int*)wvals, 0, 0, B.a,
                       B.b, B.a_eff, B.b_eff, B.inv_b, sc3[0], sc3[1], sc3[2], gp, fp, s0_d, B.closed, conv_d,
                       (const int*)wq_d + 3);
    HCHECK(hipGetLastError());
    return 0;
  }

  // every few levels: the queue's counters; once the queue is empty the run goes on with the host path (the slots handed
  // back), and *alive = 0 when nothing remains to run
  int dev_check(int* alive) {
    int wq[8];
    HCHECK(hipMemcpyAsync(wq, wq_d, sizeof(wq), hipMemcpyDeviceToHost, stream));
    HCHECK(hipStreamSynchronize(stream));
    *alive = wq[5];
    if (wq[0] >= (int)nl_all.size()) return leave_dev_queue();
    return 0;
  }

  int leave_dev_queue() {
    if (!dev_queue) return 0;
    int wq[8];
    HCHECK(hipMemcpyAsync(wq, wq_d, sizeof(wq), hipMemcpyDeviceToHost, stream));
    HCHECK(hipMemcpyAsync(slot_chain.data(), slot_chain_d, M.nbp * sizeof(int), hipMemcpyDeviceToHost, stream));
    HCHECK(hipMemcpyAsync(slot_dirty.data(), slot_dirty_d, M.nbp * sizeof(int), hipMemcpyDeviceToHost, stream));
    HCHECK(hipMemcpyAsync(slot_s0.data(), s0_d, M.nbp * sizeof(int), hipMemcpyDeviceToHost, stream));
    HCHECK(hipStreamSynchronize(stream));
    qnext = wq[0];
    dev_queue = false;
    return 0;
  }

  int level_cap() const { return win ? (nchain / std::max(1, M.nb) + 2) * (maxdepth + 4) : maxdepth; }

  int any_active(int* actv) {
    HCHECK(hipMemcpyAsync(actv, B.active, sizeof(int), hipMemcpyDeviceToHost, stream));
    HCHECK(hipStreamSynchronize(stream));
    return 0;
  }

  // coefficients in the caller's chain order; route ed3 also c_k, the closure flags and the start / projection norms
  int finish(T* a_out, T* b_out, int* mused_out, T* c_out = nullptr, int* closed_out = nullptr, double* snorm_out = nullptr,
             double* tnorm_out = nullptr, T* cpr_out = nullptr) {
    if (fused) {
      // the fused level assumes an operator symmetric on the symmetric subspace (S_RR' = S_R'R^T); when its a_k and the
      // direct <u, H u> disagree the run is void: the operator keeps the separate update pass from now on
      int bad = 0;
      HCHECK(hipMemcpyAsync(&bad, B.asym, sizeof(int), hipMemcpyDeviceToHost, stream));
      HCHECK(hipStreamSynchronize(stream));
      if (bad) {
        op->asym = true;
        fail("fused level: the operator is not symmetric on the symmetric subspace; rerun with the separate update pass");
        return RERUN_ASYM;
      }
    }
    if (win) {
      if (leave_dev_queue()) return 1;
      // the members still holding slots hand over their results, then every chain's row in the caller's order
      std::vector<int> xs, xc;
      for (int q = 0; q < M.nb; ++q)
        if (slot_chain[q] >= 0) {
          xs.push_back(q);
          xc.push_back(slot_chain[q]);
        }
      if (!xs.empty()) {
        const int n = (int)xs.size();
        HCHECK(hipMemcpyAsync(wslots, xs.data(), n * sizeof(int), hipMemcpyHostToDevice, stream));
        HCHECK(hipMemcpyAsync(wchains, xc.data(), n * sizeof(int), hipMemcpyHostToDevice, stream));
        hipLaunchKernelGGL(k_extract<T>, dim3(n), dim3(64), 0, stream, a_all, b_all, mused_all, (const T*)B.a, (const T*)B.b,
                           (const int*)B.mused, ld, (const int*)wslots, (const int*)wchains, n);
        HCHECK(hipGetLastError());
      }
      std::vector<T> ah((size_t)nchain * ld), bh((size_t)nchain * ld);
      std::vector<int> mu(nchain);
      HCHECK(hipMemcpyAsync(ah.data(), a_all, ah.size() * sizeof(T), hipMemcpyDeviceToHost, stream));
      HCHECK(hipMemcpyAsync(bh.data(), b_all, bh.size() * sizeof(T), hipMemcpyDeviceToHost, stream));
      HCHECK(hipMemcpyAsync(mu.data(), mused_all, mu.size() * sizeof(int), hipMemcpyDeviceToHost, stream));
      HCHECK(hipStreamSynchronize(stream));
      for (int q = 0; q < nchain; ++q) {
        const int c = order[q];
        std::memcpy(a_out + (size_t)c * ld, ah.data() + (size_t)q * ld, sizeof(T) * ld);
        std::memcpy(b_out + (size_t)c * ld, bh.data() + (size_t)q * ld, sizeof(T) * ld);
        mused_out[c] = mu[q];
      }
      return 0;
    }
    if (proj && cpr_out) {
      // projections in the caller's order: chain x level x target x orbital
      const size_t per = (size_t)op->ntarget * 9;
      std::vector<T> cp((size_t)ld * per * M.nbp);
      HCHECK(hipMemcpyAsync(cp.data(), B.cpr, cp.size() * sizeof(T), hipMemcpyDeviceToHost, stream));
      HCHECK(hipStreamSynchronize(stream));
      for (int q = 0; q < nchain; ++q)
        for (int k = 0; k < ld; ++k)
          for (size_t e = 0; e < per; ++e)
            cpr_out[((size_t)order[q] * ld + k) * per + e] = cp[((size_t)k * per + e) * M.nbp + q];
    }
    std::vector<T> ah((size_t)M.nbp * ld), bh((size_t)M.nbp * ld), chv(ed3 ? (size_t)M.nbp * ld : 0);
    std::vector<int> mu(M.nbp), cl(ed3 ? M.nbp : 0);
    HCHECK(hipMemcpyAsync(ah.data(), B.a, ah.size() * sizeof(T), hipMemcpyDeviceToHost, stream));
    HCHECK(hipMemcpyAsync(bh.data(), B.b, bh.size() * sizeof(T), hipMemcpyDeviceToHost, stream));
    HCHECK(hipMemcpyAsync(mu.data(), B.mused, mu.size() * sizeof(int), hipMemcpyDeviceToHost, stream));
    if (ed3) {
      HCHECK(hipMemcpyAsync(chv.data(), B.c, chv.size() * sizeof(T), hipMemcpyDeviceToHost, stream));
      HCHECK(hipMemcpyAsync(cl.data(), B.closed, cl.size() * sizeof(int), hipMemcpyDeviceToHost, stream));
    }
    HCHECK(hipStreamSynchronize(stream));
    for (int q = 0; q < nchain; ++q) {
      const int c = order[q];
      std::memcpy(a_out + (size_t)c * ld, ah.data() + (size_t)q * ld, sizeof(T) * ld);
      std::memcpy(b_out + (size_t)c * ld, bh.data() + (size_t)q * ld, sizeof(T) * ld);
      mused_out[c] = mu[q];
      if (ed3) {
        if (c_out) std::memcpy(c_out + (size_t)c * ld, chv.data() + (size_t)q * ld, sizeof(T) * ld);
        if (closed_out) closed_out[c] = cl[q];
        if (snorm_out) snorm_out[c] = snorm[q];
        if (tnorm_out) tnorm_out[c] = tnorm[q];
      }
    }
    return 0;
  }
};

template <class T>
int run_chains(asr_gpu_op* op, int nchain, const int* start_orb, const double* params, const double* pab,
               const int* frozen_sp, const double* conc, const double* z, const int* nlev, int project, int ldab, T* a_out,
               T* b_out, int* mused_out, int mode, double tiny, bool real_chain, double* elapsed_ms);

// a run of the fused level that found the operator not symmetric is repeated with the separate update pass
template <class T>
int run_chains_once(asr_gpu_op* op, int nchain, const int* start_orb, const double* params, const double* pab,
                    const int* frozen_sp, const double* conc, const double* z, const int* nlev, int project, int ldab,
                    T* a_out, T* b_out, int* mused_out, int mode, double tiny, bool real_chain, double* elapsed_ms) {
  if (nchain <= 0) return 0;
  const int nsp9 = op->nsp * 9, ncn = op->ncls * op->nsmax;
  if (mode == 1 && nchain > 1) {
    // plain form: one chain per launch sequence, host-driven levels
    for (int c = 0; c < nchain; ++c) {
      double ms = 0.0;
      int st = run_chains<T>(op, 1, start_orb + c, params + (size_t)c * 4 * nsp9, pab ? pab + (size_t)c * 2 * nsp9 : nullptr,
                             frozen_sp ? frozen_sp + c : nullptr, conc ? conc + (size_t)c * ncn : nullptr,
                             z ? z + 2 * c : nullptr, nlev + c, project, ldab, a_out + (size_t)c * ldab,
                             b_out + (size_t)c * ldab, mused_out + c, 1, tiny, real_chain, &ms);
      if (st) return st;
      if (elapsed_ms) *elapsed_ms += ms;
    }
    return 0;
  }
  ChainRun<T> R;
  if (R.setup(op, op->stream, nchain, start_orb, params, pab, frozen_sp, conc, z, nlev, project, ldab,
              mode == 1 ? 1 : op->tile_pref, tiny, real_chain, nullptr, nullptr, mode == 0))
    return 1;
  hipEvent_t e0, e1;
  HCHECK(hipEventCreate(&e0));
  HCHECK(hipEventCreate(&e1));
  HCHECK(hipEventRecord(e0, op->stream));
  const int cap = R.level_cap();
  for (int k = 0; k < cap; ++k) {
    if (R.step()) return 1;
    if (mode == 0 && R.dev_queue) {
      if (R.refill_dev()) return 1;
      if ((k & 7) == 7) {
        int alive = 1;
        if (R.dev_check(&alive)) return 1;
        if (!alive) break;
      }
    } else if (mode == 0 && (op->compact || R.win)) {
      int alive = 1;
      if (R.compact(&alive)) return 1;
      if (!alive) break;
    }
    if (mode == 1) {
      // plain form: the host reads the chain's state back after every level and stops when it has left
      if (R.flush_test()) return 1;
      int actv = 0;
      if (R.any_active(&actv)) return 1;
      if (!actv) break;
    }
  }
  if (R.flush_test()) return 1;
  HCHECK(hipEventRecord(e1, op->stream));
  HCHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  HCHECK(hipEventElapsedTime(&ms, e0, e1));
  if (elapsed_ms) *elapsed_ms = ms;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return R.finish(a_out, b_out, mused_out);
}

template <class T>
int run_chains(asr_gpu_op* op, int nchain, const int* start_orb, const double* params, const double* pab,
               const int* frozen_sp, const double* conc, const double* z, const int* nlev, int project, int ldab, T* a_out,
               T* b_out, int* mused_out, int mode, double tiny, bool real_chain, double* elapsed_ms) {
  int st = RERUN_ASYM;
  for (int pass = 0; pass < 2 && st == RERUN_ASYM; ++pass) {
    if (pass && g_timer) {  // the per-kernel timings start over with the separate update pass
      g_acc.clear();
      g_cnt.clear();
      g_levels = 0;
      g_timer->kind.clear();
    }
    st = run_chains_once<T>(op, nchain, start_orb, params, pab, frozen_sp, conc, z, nlev, project, ldab, a_out, b_out,
                            mused_out, mode, tiny, real_chain, elapsed_ms);
  }
  return st;
}

// Chains on several operators at once (one group per operator: the sublattices of a multi-sublattice alloy).  Every group
// runs on its own stream so that small operators fill the GPU together; members of a group are sorted by depth and tiles
// whose members have all left are skipped.  Per-chain arrays are concatenated over the groups in order; conc (optional)
// holds per chain the ncls x nsmax concentrations of its own operator.  mode 0: host-issued levels on the group streams;
// mode 2: the whole run captured into one HIP graph and launched once; mode 1: the plain form, one chain per launch sequence.
template <class T>
int run_multi_once(int ngroup, asr_gpu_op* const* ops, const int* ngch, const int* start_orb, const double* params,
                   const double* pab, const int* frozen_sp, const double* conc, const double* z, const int* nlev,
                   int project, int ldab, T* a, T* b, int* mused, int mode, double tiny, bool real_chain,
                   double* elapsed_ms) {
  if (elapsed_ms) *elapsed_ms = 0.0;
  std::vector<size_t> off_c(ngroup + 1, 0), off_p(ngroup + 1, 0), off_pa(ngroup + 1, 0), off_cc(ngroup + 1, 0);
  for (int g = 0; g < ngroup; ++g) {
    off_c[g + 1] = off_c[g] + ngch[g];
    off_p[g + 1] = off_p[g] + (size_t)ngch[g] * 4 * ops[g]->nsp * 9;
    off_pa[g + 1] = off_pa[g] + (size_t)ngch[g] * 2 * ops[g]->nsp * 9;
    off_cc[g + 1] = off_cc[g] + (size_t)ngch[g] * ops[g]->ncls * ops[g]->nsmax;
  }
  if (mode == 1) {
    for (int g = 0; g < ngroup; ++g) {
      double ms = 0.0;
      const size_t c0 = off_c[g];
      if (run_chains<T>(ops[g], ngch[g], start_orb + c0, params + off_p[g], pab ? pab + off_pa[g] : nullptr,
                        frozen_sp ? frozen_sp + c0 : nullptr, conc ? conc + off_cc[g] : nullptr, z ? z + 2 * c0 : nullptr,
                        nlev + c0, project, ldab, a + c0 * ldab, b + c0 * ldab, mused + c0, 1, tiny, real_chain, &ms))
        return 1;
      if (elapsed_ms) *elapsed_ms += ms;
    }
    return 0;
  }
  // the adaptive depth with the bp tail decides on the host between levels: no graph capture then (the other tails are
  // tested on the device, k_depth_test)
  for (int g = 0; g < ngroup; ++g)
    if (mode == 2 && ops[g]->depth_tol > 0.0 && ops[g]->depth_tail == 1 && std::is_same<T, cplx>::value) mode = 0;
  std::vector<hipStream_t> st(ngroup, nullptr);
  std::vector<ChainRun<T>> runs(ngroup);
  auto cleanup = [&]() {
    runs.clear();  // the runs synchronise on their streams when they release their buffers: before the streams go
    for (auto s2 : st)
      if (s2) (void)hipStreamDestroy(s2);
  };
  for (int g = 0; g < ngroup; ++g) {
    if (hipStreamCreateWithFlags(&st[g], hipStreamNonBlocking) != hipSuccess) { cleanup(); return fail("stream create"); }
    const size_t c0 = off_c[g];
    if (ngch[g] <= 0) continue;
    if (runs[g].setup(ops[g], st[g], ngch[g], start_orb + c0, params + off_p[g], pab ? pab + off_pa[g] : nullptr,
                      frozen_sp ? frozen_sp + c0 : nullptr, conc ? conc + off_cc[g] : nullptr, z ? z + 2 * c0 : nullptr,
                      nlev + c0, project, ldab, ops[g]->tile_pref, tiny, real_chain)) {
      cleanup();
      return 1;
    }
  }
  hipStream_t s0 = st[0];
  hipEvent_t e0, e1, fork;
  std::vector<hipEvent_t> join(ngroup);
  HCHECK(hipEventCreate(&e0));
  HCHECK(hipEventCreate(&e1));
  HCHECK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
  for (auto& e : join) HCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  int maxd = 0;
  for (int g = 0; g < ngroup; ++g) maxd = std::max(maxd, ngch[g] > 0 ? runs[g].maxdepth : 0);
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int status = 0;
  auto issue = [&]() -> int {
    // fork every group stream off s0, levels interleaved over the groups, join back into s0
    if (hipEventRecord(fork, s0) != hipSuccess) return fail("event record");
    for (int g = 1; g < ngroup; ++g)
      if (hipStreamWaitEvent(st[g], fork, 0) != hipSuccess) return fail("stream wait");
    for (int k = 0; k < maxd; ++k)
      for (int g = 0; g < ngroup; ++g) {
        if (ngch[g] > 0 && k < runs[g].maxdepth && runs[g].step()) return 1;
        // host-issued levels: member compaction after the levels of the depth test (a group whose members have all
        // left issues no more levels)
        if (mode == 0 && ngch[g] > 0 && k < runs[g].maxdepth && ops[g]->compact) {
          int alive = 1;
          if (runs[g].compact(&alive)) return 1;
          if (!alive) runs[g].maxdepth = k + 1;
        }
      }
    for (int g = 0; g < ngroup; ++g)
      if (ngch[g] > 0 && runs[g].flush_test()) return 1;
    for (int g = 1; g < ngroup; ++g) {
      if (hipEventRecord(join[g], st[g]) != hipSuccess) return fail("event record");
      if (hipStreamWaitEvent(s0, join[g], 0) != hipSuccess) return fail("stream wait");
    }
    return 0;
  };
  if (mode == 2) {
    HCHECK(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal));
    status = issue();
    hipError_t ce = hipStreamEndCapture(s0, &graph);
    if (!status && ce != hipSuccess) status = fail(std::string("graph capture: ") + hipGetErrorString(ce));
    if (!status && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) status = fail("graph instantiate");
    if (!status) {
      HCHECK(hipEventRecord(e0, s0));
      if (hipGraphLaunch(exec, s0) != hipSuccess) status = fail("graph launch");
      HCHECK(hipEventRecord(e1, s0));
    }
  } else {
    HCHECK(hipEventRecord(e0, s0));
    status = issue();
    HCHECK(hipEventRecord(e1, s0));
  }
  if (!status) {
    HCHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    HCHECK(hipEventElapsedTime(&ms, e0, e1));
    if (elapsed_ms) *elapsed_ms = ms;
    for (int g = 0; g < ngroup && !status; ++g) {
      const size_t c0 = off_c[g];
      if (ngch[g] > 0) status = runs[g].finish(a + c0 * ldab, b + c0 * ldab, mused + c0);
    }
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  for (hipEvent_t e : {e0, e1, fork}) (void)hipEventDestroy(e);
  for (auto& e : join) (void)hipEventDestroy(e);
  cleanup();
  return status;
}

template <class T>
int run_multi(int ngroup, asr_gpu_op* const* ops, const int* ngch, const int* start_orb, const double* params,
              const double* pab, const int* frozen_sp, const double* conc, const double* z, const int* nlev, int project,
              int ldab, T* a, T* b, int* mused, int mode, double tiny, bool real_chain, double* elapsed_ms) {
  int st = RERUN_ASYM;
  for (int pass = 0; pass < 2 && st == RERUN_ASYM; ++pass)
    st = run_multi_once<T>(ngroup, ops, ngch, start_orb, params, pab, frozen_sp, conc, z, nlev, project, ldab, a, b, mused,
                           mode, tiny, real_chain, elapsed_ms);
  return st;
}

// The chains of two-chain blocks on one operator (the members of asr_gpu_ed3_blocks and asr_gpu_epi_blocks, two per
// block): mode 0 host-issued levels, 1 the plain form (one chain per launch sequence, host read-back per level, as chain),
// 2 the whole run as one HIP graph (with the adaptive depth on the device; the bp tail's test runs on the host, so it
// falls back to mode 0).
int run_blocks(asr_gpu_op* op, int nmem, const int* so, const double* params, const double* pab, const int* fz,
               const double* conc, const double* z, const int* nlev, const BlockIn& e3, int project, int ldab, cplx* a, cplx* b,
               cplx* c, int* mused, int* closed, double* snorm, double* tnorm, int mode, double* elapsed_ms);
int run_blocks_once(asr_gpu_op* op, int nmem, const int* so, const double* params, const double* pab, const int* fz,
                    const double* conc, const double* z, const int* nlev, const BlockIn& e3, int project, int ldab, cplx* a,
                    cplx* b, cplx* c, int* mused, int* closed, double* snorm, double* tnorm, int mode, double* elapsed_ms) {
  const int nsp9 = op->nsp * 9, ncn = op->ncls * op->nsmax;
  *elapsed_ms = 0.0;
  if (mode == 2 && op->depth_tol > 0.0 && op->depth_tail == 1) mode = 0;
  if (mode == 1 && nmem > 1) {
    for (int q = 0; q < nmem; ++q) {
      const BlockIn e1{e3.pex ? e3.pex + (size_t)q * 4 * nsp9 : nullptr, e3.oc + q, e3.side + q,
                       e3.v9 ? e3.v9 + (size_t)q * 9 : nullptr, e3.chim ? e3.chim + q : nullptr,
                       e3.fsp2 ? e3.fsp2 + q : nullptr};
      double ms = 0.0;
      if (run_blocks(op, 1, so + q, params + (size_t)q * 4 * nsp9, pab + (size_t)q * 2 * nsp9, fz ? fz + q : nullptr,
                  conc ? conc + (size_t)q * ncn : nullptr, z + 2 * q, nlev + q, e1, project, ldab, a + (size_t)q * ldab,
                  b + (size_t)q * ldab, c + (size_t)q * ldab, mused + q, closed + q, snorm + q, tnorm + q, 1, &ms))
        return 1;
      *elapsed_ms += ms;
    }
    return 0;
  }
  ChainRun<cplx> R;
  if (R.setup(op, op->stream, nmem, so, params, pab, fz, conc, z, nlev, project, ldab, mode == 1 ? 1 : op->tile_pref, 1.0e-12,
              false, &e3))
    return 1;
  hipEvent_t e0, e1;
  HCHECK(hipEventCreate(&e0));
  HCHECK(hipEventCreate(&e1));
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int status = 0;
  if (mode == 2) {
    HCHECK(hipStreamBeginCapture(op->stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < R.maxdepth && !status; ++k) status = R.step();
    if (!status) status = R.flush_test();
    hipError_t ce = hipStreamEndCapture(op->stream, &graph);
    if (!status && ce != hipSuccess) status = fail(std::string("graph capture: ") + hipGetErrorString(ce));
    if (!status && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) status = fail("graph instantiate");
    if (!status) {
      HCHECK(hipEventRecord(e0, op->stream));
      if (hipGraphLaunch(exec, op->stream) != hipSuccess) status = fail("graph launch");
      HCHECK(hipEventRecord(e1, op->stream));
    }
  } else {
    HCHECK(hipEventRecord(e0, op->stream));
    for (int k = 0; k < R.maxdepth && !status; ++k) {
      status = R.step();
      if (!status && mode == 0 && op->compact) {
        int alive = 1;
        status = R.compact(&alive);
        if (!status && !alive) break;
      }
      if (!status && mode == 1) {
        status = R.flush_test();
        int actv = 0;
        if (!status) status = R.any_active(&actv);
        if (!status && !actv) break;
      }
    }
    if (!status) status = R.flush_test();
    HCHECK(hipEventRecord(e1, op->stream));
  }
  if (!status) {
    HCHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    HCHECK(hipEventElapsedTime(&ms, e0, e1));
    *elapsed_ms = ms;
    status = R.finish(a, b, mused, c, closed, snorm, tnorm);
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return status;
}

// a run whose fused level found the operator not symmetric on the symmetric subspace is repeated with the separate
// update pass (the operator keeps that form)
int run_blocks(asr_gpu_op* op, int nmem, const int* so, const double* params, const double* pab, const int* fz,
               const double* conc, const double* z, const int* nlev, const BlockIn& e3, int project, int ldab, cplx* a, cplx* b,
               cplx* c, int* mused, int* closed, double* snorm, double* tnorm, int mode, double* elapsed_ms) {
  int st = RERUN_ASYM;
  for (int pass = 0; pass < 2 && st == RERUN_ASYM; ++pass)
    st = run_blocks_once(op, nmem, so, params, pab, fz, conc, z, nlev, e3, project, ldab, a, b, c, mused, closed, snorm,
                         tnorm, mode, elapsed_ms);
  return st;
}

// Projected chains of route edc on one operator (the chains of asr_gpu_chains_edc_proj): mode 0 host-issued levels, 1 the
// plain form (one chain per launch sequence), 2 the whole run as one HIP graph.  cpr: nchain x ldab x ntarget x 9.
int run_proj(asr_gpu_op* op, int nchain, const int* so, const double* params, const double* pab, const int* fz,
             const double* conc, const double* z, const int* nlev, const double* tw, int project, int ldab, cplx* a, cplx* b,
             int* mused, cplx* cpr, int mode, double* elapsed_ms);

int run_proj_once(asr_gpu_op* op, int nchain, const int* so, const double* params, const double* pab, const int* fz,
                  const double* conc, const double* z, const int* nlev, const double* tw, int project, int ldab, cplx* a,
                  cplx* b, int* mused, cplx* cpr, int mode, double* elapsed_ms) {
  const int nsp9 = op->nsp * 9, ncn = op->ncls * op->nsmax, ntw = op->ntarget * op->nsmax * 9;
  *elapsed_ms = 0.0;
  if (mode == 1 && nchain > 1) {
    for (int q = 0; q < nchain; ++q) {
      double ms = 0.0;
      if (run_proj(op, 1, so + q, params + (size_t)q * 4 * nsp9, pab + (size_t)q * 2 * nsp9, fz ? fz + q : nullptr,
                   conc ? conc + (size_t)q * ncn : nullptr, z + 2 * q, nlev + q, tw + (size_t)q * ntw, project, ldab,
                   a + (size_t)q * ldab, b + (size_t)q * ldab, mused + q, cpr + (size_t)q * ldab * op->ntarget * 9, 1, &ms))
        return 1;
      *elapsed_ms += ms;
    }
    return 0;
  }
  ChainRun<cplx> R;
  if (R.setup(op, op->stream, nchain, so, params, pab, fz, conc, z, nlev, project, ldab, mode == 1 ? 1 : op->tile_pref,
              1.0e-12, false, nullptr, tw))
    return 1;
  hipEvent_t e0, e1;
  HCHECK(hipEventCreate(&e0));
  HCHECK(hipEventCreate(&e1));
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int status = 0;
  if (mode == 2) {
    HCHECK(hipStreamBeginCapture(op->stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < R.maxdepth && !status; ++k) status = R.step();
    hipError_t ce = hipStreamEndCapture(op->stream, &graph);
    if (!status && ce != hipSuccess) status = fail(std::string("graph capture: ") + hipGetErrorString(ce));
    if (!status && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) status = fail("graph instantiate");
    if (!status) {
      HCHECK(hipEventRecord(e0, op->stream));
      if (hipGraphLaunch(exec, op->stream) != hipSuccess) status = fail("graph launch");
      HCHECK(hipEventRecord(e1, op->stream));
    }
  } else {
    HCHECK(hipEventRecord(e0, op->stream));
    for (int k = 0; k < R.maxdepth && !status; ++k) {
      status = R.step();
      if (!status && mode == 1) {
        int actv = 0;
        status = R.any_active(&actv);
        if (!status && !actv) break;
      }
    }
    HCHECK(hipEventRecord(e1, op->stream));
  }
  if (!status) {
    HCHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    HCHECK(hipEventElapsedTime(&ms, e0, e1));
    *elapsed_ms = ms;
    status = R.finish(a, b, mused, nullptr, nullptr, nullptr, nullptr, cpr);
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return status;
}

int run_proj(asr_gpu_op* op, int nchain, const int* so, const double* params, const double* pab, const int* fz,
             const double* conc, const double* z, const int* nlev, const double* tw, int project, int ldab, cplx* a, cplx* b,
             int* mused, cplx* cpr, int mode, double* elapsed_ms) {
  int st = RERUN_ASYM;
  for (int pass = 0; pass < 2 && st == RERUN_ASYM; ++pass)
    st = run_proj_once(op, nchain, so, params, pab, fz, conc, z, nlev, tw, project, ldab, a, b, mused, cpr, mode, elapsed_ms);
  return st;
}

// Adjoint of the real Lanczos recursion (the reference has none; the forward is lanczos, src/reference/asr_recursion.f90:17-58)
// for chains of equal depth nlev on H = c + d S d (o = E_nu = 0).  Forward: the batched chains, (u_k, u_{k-1}) checkpointed
// every ckpt levels.  Reverse, segment by segment: the levels of the segment are recomputed from its checkpoint (keeping
// u_k, w_k = H u_k and S D u~_k), then for k from the segment's end down to its start
//   b_bar_k = dL/db_k - <r_bar_{k+1}, u_k> - <u_bar_{k+1}, u_{k+1}> / b_k,   r_bar_k = u_bar_{k+1} / b_k + b_bar_k u_{k+1}
//   a_bar_k = dL/da_k - <r_bar_k, u_k>,   w_bar_k = r_bar_k + a_bar_k u_k
//   u_bar_k = -b_k r_bar_{k+1} - a_k r_bar_k + a_bar_k w_k + H w_bar_k,   grad += <w_bar_k, (dH/dlambda) u_k>
// (r_k = w_k - a_k u_k - b_{k-1} u_{k-1}, u_{k+1} = r_k / b_k; u_bar_nlev = 0).  dL/da, dL/db: nchain x nlev.
// grad: nchain x 2 x nsp x 3 (dL/dc, dL/dd per species and l shell).  Cost: one forward, one recomputed forward and one
// operator application per level on the adjoint vector.
int run_adjoint(asr_gpu_op* op, int nchain, const int* start_orb, const double* params, const int* frozen_sp,
                const double* conc, int nlev, int ckpt, const double* La, const double* Lb, double* a_out, double* b_out,
                double* grad, double* fwd_ms, double* adj_ms) {
  const int nsp = op->nsp;
  if (nsp > 5) return fail("lanczos_adjoint: at most 5 species");
  if (ckpt < 1) ckpt = 8;
  for (int c = 0; c < nchain; ++c)
    for (int q = 2 * nsp * 9; q < 4 * nsp * 9; ++q)
      if (params[(size_t)c * 4 * nsp * 9 + q] != 0.0)
        return fail("lanczos_adjoint: the adjoint is written for H = c + d S d (o = E_nu = 0)");
  std::vector<int> nl(nchain, nlev);
  ChainRun<double> R;
  if (R.setup(op, op->stream, nchain, start_orb, params, nullptr, frozen_sp, conc, nullptr, nl.data(), 0, nlev,
              op->tile_pref, 1e-300, true))
    return 1;
  R.M.uniform_d = false;  // every application materialises S D x~ (the general path)
  hipStream_t st = R.stream;
  const size_t nv = R.nv;
  const int nseg = (nlev + ckpt - 1) / ckpt;
  const MemDev md0 = R.md;
  std::vector<double*> bufs;
  auto vec = [&](double** p) -> int {
    if (dalloc(p, nv)) return 1;
    bufs.push_back(*p);
    return hipMemsetAsync(*p, 0, nv * 8, st) != hipSuccess;
  };
  struct Free {
    std::vector<double*>* b;
    ~Free() {
      for (double* p : *b) (void)hipFree(p);
    }
  } freer{&bufs};
  std::vector<double*> CK(2 * (nseg - 1)), U(ckpt + 2), W(ckpt), T2(ckpt);  // no checkpoint for the last segment
  for (auto& p : CK) if (vec(&p)) return 1;
  for (auto& p : U) if (vec(&p)) return 1;
  for (auto& p : W) if (vec(&p)) return 1;
  for (auto& p : T2) if (vec(&p)) return 1;
  double *rbar, *rprev, *hr, *t2r, *ubar, *unext;
  for (double** p : {&rbar, &rprev, &hr, &t2r, &ubar, &unext}) if (vec(p)) return 1;
  const dim3 grid = R.grid;
  const int nbx = grid.x, mt = R.M.mt, nbp = R.M.nbp;
  const size_t plane = (size_t)grid.x * grid.y * mt;
  const int nq = 2 * nsp * 3;
  // the gradient pass walks the node blocks with at most 1024 blocks per member tile (a multiple of the 8 XCDs)
  const int gxg = std::min(((nbx + 7) / 8) * 8, 1024);
  const dim3 ggrid(gxg, grid.y);
  const bool gbuf = buffer_ok(op, R.md, sizeof(double));  // own rows of the gradient pass as 32-bit buffer loads
  const size_t gplane = (size_t)gxg * grid.y * mt;
  double *P1, *P2, *P3, *PG, *coef, *gacc, *dLa, *dLb;
  for (double** p : {&P1, &P2, &P3}) {
    if (dalloc(p, plane)) return 1;
    bufs.push_back(*p);
  }
  if (dalloc(&PG, gplane * nq) || dalloc(&coef, (size_t)12 * nbp) || dalloc(&gacc, (size_t)nbp * nq) ||
      dalloc(&dLa, (size_t)nbp * nlev) || dalloc(&dLb, (size_t)nbp * nlev))
    return 1;
  for (double* p : {PG, coef, gacc, dLa, dLb}) bufs.push_back(p);
  double *a_eff = coef, *b_prev = coef + nbp, *inv_b = coef + 2 * nbp, *nega = coef + 3 * nbp, *negb = coef + 4 * nbp,
         *alpha = coef + 5 * nbp, *beta = coef + 6 * nbp, *abar = coef + 7 * nbp, *cbb = coef + 8 * nbp,
         *abar2 = coef + 9 * nbp;
  HCHECK(hipMemsetAsync(coef, 0, (size_t)12 * nbp * 8, st));
  HCHECK(hipMemsetAsync(gacc, 0, (size_t)nbp * nq * 8, st));
  HCHECK(hipMemsetAsync(PG, 0, gplane * nq * 8, st));
  HCHECK(hipMemsetAsync(P1, 0, plane * 8, st));
  {
    // loss derivatives in the sorted member order
    std::vector<double> la((size_t)nbp * nlev, 0.0), lb((size_t)nbp * nlev, 0.0);
    for (int q = 0; q < nchain; ++q) {
      std::memcpy(&la[(size_t)q * nlev], La + (size_t)R.order[q] * nlev, 8 * nlev);
      std::memcpy(&lb[(size_t)q * nlev], Lb + (size_t)R.order[q] * nlev, 8 * nlev);
    }
    HCHECK(hipMemcpyAsync(dLa, la.data(), la.size() * 8, hipMemcpyHostToDevice, st));
    HCHECK(hipMemcpyAsync(dLb, lb.data(), lb.size() * 8, hipMemcpyHostToDevice, st));
  }
  int* spc_d = nullptr;
  if (dalloc(&spc_d, (size_t)op->ncls * op->nsmax)) return 1;
  HCHECK(hipMemcpy(spc_d, op->spc_h.data(), op->spc_h.size() * 4, hipMemcpyHostToDevice));
  struct FreeI {
    int* p;
    ~FreeI() { (void)hipFree(p); }
  } freei{spc_d};
  const OpDev od = dev_of(op);
  auto lincomb = [&](double* out, int nt, const double* v0, const double* c0, const double* v1, const double* c1,
                     const double* v2, const double* c2, const double* v3, const double* c3, int np, const double* p0,
                     double* q0, const double* p1, double* q1) -> int {
    hipLaunchKernelGGL(k_lincomb, grid, dim3(BLOCK), 0, st, od, md0, out, nt, v0, c0, v1, c1, v2, c2, v3, c3, np, p0, q0, p1,
                       q1);
    HCHECK(hipGetLastError());
    return 0;
  };
  // the recomputation and the u_bar combination ride on the operator's last pass (merged_last_pass), else separate passes
  const bool fuse = op->level_form == 0 && merged_last_pass(op, R.M, md0, sizeof(double), false);
  hipEvent_t e0, e1, e2;
  HCHECK(hipEventCreate(&e0));
  HCHECK(hipEventCreate(&e1));
  HCHECK(hipEventCreate(&e2));
  // ---- forward with checkpoints; the last segment keeps its level vectors (u_k, H u_k, S D u~_k) in the segment buffers,
  // so the reverse sweep starts without recomputing it
  const int klast = (nseg - 1) * ckpt;
  HCHECK(hipEventRecord(e0, st));
  for (int k = 0; k < nlev; ++k) {
    if (k >= klast) {
      const int j = k - klast;
      if (j == 0) {
        HCHECK(hipMemcpyAsync(U[1], R.B.u, nv * 8, hipMemcpyDeviceToDevice, st));
        HCHECK(hipMemcpyAsync(U[0], R.B.up, nv * 8, hipMemcpyDeviceToDevice, st));
      }
      double* t2s = R.B.t2;
      R.B.t2 = T2[j];
      if (R.step_apply()) return 1;
      R.B.t2 = t2s;
      HCHECK(hipMemcpyAsync(W[j], R.B.w, nv * 8, hipMemcpyDeviceToDevice, st));  // H u_k before the update
      if (R.step_update()) return 1;
      // u_nlev is zero in the forward (a member leaves at its last level); the reverse takes it from the recorded b below
      if (k + 1 < nlev) HCHECK(hipMemcpyAsync(U[2 + j], R.B.u, nv * 8, hipMemcpyDeviceToDevice, st));
      continue;
    }
    if (k % ckpt == 0) {
      HCHECK(hipMemcpyAsync(CK[2 * (k / ckpt)], R.B.u, nv * 8, hipMemcpyDeviceToDevice, st));
      HCHECK(hipMemcpyAsync(CK[2 * (k / ckpt) + 1], R.B.up, nv * 8, hipMemcpyDeviceToDevice, st));
    }
    if (R.step()) return 1;
  }
  HCHECK(hipEventRecord(e1, st));
  // ---- reverse: u_bar_nlev = 0, r_bar_nlev = 0, <u_bar_nlev, u_nlev> = 0 (P1 cleared)
  for (int seg = nseg - 1; seg >= 0; --seg) {
    const int k0 = seg * ckpt, k1 = std::min(nlev, k0 + ckpt) - 1;
    // U[0] = u_{k0 - 1}, U[1 + j] = u_{k0 + j} (the last segment's are in place from the forward)
    if (seg != nseg - 1) {
      HCHECK(hipMemcpyAsync(U[1], CK[2 * seg], nv * 8, hipMemcpyDeviceToDevice, st));
      HCHECK(hipMemcpyAsync(U[0], CK[2 * seg + 1], nv * 8, hipMemcpyDeviceToDevice, st));
    } else {
      // u_nlev = (H u_k1 - a_k1 u_k1 - b_k1-1 u_k1-1) / b_k1 as a recomputed level would give it
      const int j = k1 - k0;
      hipLaunchKernelGGL(k_adj_coef, dim3((nbp + 63) / 64), dim3(64), 0, st, R.M.nb, nlev, k1, R.B.a, R.B.b, a_eff, b_prev,
                         inv_b, nega, negb);
      hipLaunchKernelGGL(k_recompute, grid, dim3(BLOCK), 0, st, od, md0, U[2 + j], W[j], U[1 + j], U[j], a_eff, b_prev, inv_b);
      HCHECK(hipGetLastError());
    }
    for (int k = k0; k <= k1 && seg != nseg - 1; ++k) {
      const int j = k - k0;
      hipLaunchKernelGGL(k_adj_coef, dim3((nbp + 63) / 64), dim3(64), 0, st, R.M.nb, nlev, k, R.B.a, R.B.b, a_eff, b_prev,
                         inv_b, nega, negb);
      if (fuse) {
        // u_{k+1} from the operator's output in the same pass
        RealEpi e;
        e.out3 = U[2 + j]; e.u3 = U[1 + j]; e.up3 = U[j]; e.ra = a_eff; e.rb = b_prev; e.rib = inv_b;
        if (apply_chain_op<double>(op, R.M, md0, U[1 + j], W[j], R.B.t1, T2[j], R.B.h1, false, nullptr, nullptr, st, &e))
          return 1;
      } else {
        if (apply_chain_op<double>(op, R.M, md0, U[1 + j], W[j], R.B.t1, T2[j], R.B.h1, false, nullptr, nullptr, st))
          return 1;
        hipLaunchKernelGGL(k_recompute, grid, dim3(BLOCK), 0, st, od, md0, U[2 + j], W[j], U[1 + j], U[j], a_eff, b_prev,
                           inv_b);
        HCHECK(hipGetLastError());
      }
    }
    for (int k = k1; k >= k0; --k) {
      const int j = k - k0;
      hipLaunchKernelGGL(k_adj_coef, dim3((nbp + 63) / 64), dim3(64), 0, st, R.M.nb, nlev, k, R.B.a, R.B.b, a_eff, b_prev,
                         inv_b, nega, negb);
      // b_bar_k from d1 = <u_bar_{k+1}, u_{k+1}> (partials of the previous step), then r_bar_k with <r_bar, u_k>,
      // <r_bar, u_{k-1}>, then a_bar_k
      hipLaunchKernelGGL(k_adj_fin1, dim3(R.M.nb), dim3(BLOCK), 0, st, R.M.nb, nlev, k, P1, nbx, mt, R.B.b, dLb, cbb, alpha,
                         beta);
      if (fuse) {
        // r_bar with its dots and the operator's first pass on it (t1) in one sweep
        hipLaunchKernelGGL(k_adj_dpass, grid, dim3(BLOCK), 0, st, od, md0, unext, U[2 + j], alpha, beta, U[1 + j], U[j], rbar,
                           R.B.t1, P2, P3);
        HCHECK(hipGetLastError());
      } else if (lincomb(rbar, 2, unext, alpha, U[2 + j], beta, nullptr, nullptr, nullptr, nullptr, 2, U[1 + j], P2, U[j], P3)) {
        return 1;
      }
      hipLaunchKernelGGL(k_adj_fin2, dim3(R.M.nb), dim3(BLOCK), 0, st, R.M.nb, nlev, k, P2, P3, nbx, mt, dLa, abar, abar2,
                         cbb);
      // w_bar = r_bar + a_bar u_k enters linearly: H w_bar = H r_bar + a_bar w_k, S D w~_bar = S D r~_bar + a_bar S D u~_k
      // fused: the operator pass on r_bar writes u_bar_k = -b_k r_bar_{k+1} - a_k r_bar_k + 2 a_bar_k w_k + H r_bar_k
      // with <u_bar_k, u_k> (below) at once
      if (fuse) {
        RealEpi e;
        e.t1_ready = true;
        e.lcn = 3;
        e.lcv[0] = rprev; e.lcc[0] = negb;
        e.lcv[1] = rbar; e.lcc[1] = nega;
        e.lcv[2] = W[j]; e.lcc[2] = abar2;
        if (apply_chain_op<double>(op, R.M, md0, rbar, ubar, R.B.t1, t2r, R.B.h1, false, U[1 + j], P1, st, &e)) return 1;
      } else if (apply_chain_op<double>(op, R.M, md0, rbar, hr, R.B.t1, t2r, R.B.h1, false, nullptr, nullptr, st)) {
        return 1;
      }
#define ASR_GRAD(N) do { if (gbuf) hipLaunchKernelGGL((k_adj_grad<N, true>), ggrid, dim3(BLOCK), 0, st, od, md0, U[1 + j], rbar, abar, \
                                                    T2[j], t2r, spc_d, PG, nbx); \
                            else hipLaunchKernelGGL((k_adj_grad<N, false>), ggrid, dim3(BLOCK), 0, st, od, md0, U[1 + j], rbar, \
                                                    abar, T2[j], t2r, spc_d, PG, nbx); } while (0)
      switch (nsp) {
        case 1: ASR_GRAD(1); break;
        case 2: ASR_GRAD(2); break;
        case 3: ASR_GRAD(3); break;
        case 4: ASR_GRAD(4); break;
        default: ASR_GRAD(5); break;
      }
#undef ASR_GRAD
      // u_bar_k = -b_k r_bar_{k+1} - a_k r_bar_k + 2 a_bar_k w_k + H r_bar_k, with <u_bar_k, u_k> for the next level's b_bar
      if (!fuse && lincomb(ubar, 4, rprev, negb, rbar, nega, W[j], abar2, hr, nullptr, 1, U[1 + j], P1, nullptr, nullptr))
        return 1;
      std::swap(unext, ubar);
      std::swap(rprev, rbar);
    }
  }
  hipLaunchKernelGGL(k_adj_fin_grad, dim3(R.M.nb), dim3(BLOCK), 0, st, R.M.nb, nq, PG, gxg, mt, gplane, gacc);
  HCHECK(hipEventRecord(e2, st));
  HCHECK(hipGetLastError());
  HCHECK(hipEventSynchronize(e2));
  float f1 = 0.f, f2 = 0.f;
  HCHECK(hipEventElapsedTime(&f1, e0, e1));
  HCHECK(hipEventElapsedTime(&f2, e1, e2));
  if (fwd_ms) *fwd_ms = f1;
  if (adj_ms) *adj_ms = f2;
  for (hipEvent_t e : {e0, e1, e2}) (void)hipEventDestroy(e);
  std::vector<double> g((size_t)nbp * nq);
  HCHECK(hipMemcpy(g.data(), gacc, g.size() * 8, hipMemcpyDeviceToHost));
  for (int q = 0; q < nchain; ++q) std::memcpy(grad + (size_t)R.order[q] * nq, &g[(size_t)q * nq], 8 * nq);
  std::vector<int> mu(nchain);
  return R.finish(a_out, b_out, mu.data());
}

// host layout [member][dim] <-> device layout [(node*9 + l) * nbp + member]
template <class T>
void to_dev_layout(const T* host, std::vector<T>& dev, int nb, int nbp, size_t dim) {
  dev.assign(dim * nbp, T());
  std::memset(dev.data(), 0, dev.size() * sizeof(T));
  for (int m = 0; m < nb; ++m)
    for (size_t q = 0; q < dim; ++q) dev[q * nbp + m] = host[(size_t)m * dim + q];
}
template <class T>
void from_dev_layout(const std::vector<T>& dev, T* host, int nb, int nbp, size_t dim) {
  for (int m = 0; m < nb; ++m)
    for (size_t q = 0; q < dim; ++q) host[(size_t)m * dim + q] = dev[q * nbp + m];
}

}  // namespace

// ================================================================================================ C ABI
extern "C" {

const char* asr_gpu_last_error(void) { return g_err.c_str(); }

int asr_gpu_device_count(int* n) {
  HCHECK(hipGetDeviceCount(n));
  return 0;
}

int asr_gpu_op_create(int nn, int nslot, int nop, int nsmax, int ncls, int nsp, int ntype, asr_gpu_op** out) {
  *out = nullptr;
  if (nn <= 0 || nslot < 0 || nop <= 0 || nop > 64 || nsmax <= 0 || nsmax > 16 || ncls <= 0 || ncls > 255 || nsp <= 0 ||
      ntype <= 0 || ntype > 255)
    return fail("asr_gpu_op_create: dimensions out of range (nop <= 64, nsmax <= 16, ncls, ntype <= 255)");
  if ((size_t)nn + 1 >= (1u << 26)) return fail("asr_gpu_op_create: 2^26 nodes or more");
  auto* op = new asr_gpu_op();
  op->nn = nn; op->nslot = nslot; op->nop = nop; op->nsmax = nsmax; op->ncls = ncls; op->nsp = nsp; op->ntype = ntype;
  op->nstab = 0;
  hipError_t e = hipStreamCreateWithFlags(&op->stream, hipStreamNonBlocking);
  if (e != hipSuccess) { delete op; return fail(std::string("hipStreamCreate: ") + hipGetErrorString(e)); }
  *out = op;
  return 0;
}

int asr_gpu_op_set_tables(asr_gpu_op* op, const int* hopr, const int* hopg, const int* flipr, const int* flipg, const int* lst,
                          const int* ncl, const int* ntyp, const signed char* fz, const double* sq, const int* stab_ptr,
                          const int* stab, int nstab) {
  if (!op) return fail("asr_gpu_op_set_tables: null operator");
  const int nn = op->nn, ns = op->nslot, nsm = op->nsmax;
  std::vector<uint32_t> hop((size_t)nn * ns), flip((size_t)nn * nsm), info(nn);
  std::vector<int> plist;  // the nodes with a flip partner other than themselves
  for (size_t q = 0; q < hop.size(); ++q) {
    if (hopr[q] >= nn || (hopr[q] >= 0 && (hopg[q] < 0 || hopg[q] >= op->nop))) return fail("set_tables: hop link out of range");
    // a missing link points at the padding node nn, whose amplitudes are zero in every vector
    hop[q] = hopr[q] < 0 ? ((uint32_t)(nn + 1) << 6) : ((uint32_t)(hopr[q] + 1) << 6) | (uint32_t)hopg[q];
  }
  for (size_t q = 0; q < flip.size(); ++q) {
    if (flipr[q] >= nn || (flipr[q] >= 0 && (flipg[q] < 0 || flipg[q] >= op->nop)))
      return fail("set_tables: flip link out of range");
    flip[q] = flipr[q] < 0 ? 0u : ((uint32_t)(flipr[q] + 1) << 6) | (uint32_t)flipg[q];
  }
  op->slot_flippable[0] = op->slot_flippable[1] = false;
  for (int i = 0; i < nn; ++i) {
    if (lst[i] < 0 || lst[i] >= nsm || ncl[i] < 0 || ncl[i] >= op->ncls || ntyp[i] < 0 || ntyp[i] >= op->ntype)
      return fail("set_tables: node state, class or type out of range");
    if (fz && (fz[i] < 0 || fz[i] > 2)) return fail("set_tables: frozen-site slot 0 (none), 1 (start site) or 2 (pair partner)");
    // a site that can be frozen must have been non-flippable: its nodes keep the reference state, no flip partner
    if (fz && fz[i]) {
      bool flips = lst[i] != 0;
      for (int k = 0; k < nsm; ++k) flips |= (k != lst[i] && flipr[(size_t)i * nsm + k] >= 0);
      if (flips) op->slot_flippable[fz[i] - 1] = true;
    }
    info[i] = (uint32_t)lst[i] | ((uint32_t)ncl[i] << 4) | ((uint32_t)ntyp[i] << 12) | ((uint32_t)(fz ? fz[i] : 0) << 20);
    bool partners = false;
    for (int k = 0; k < nsm; ++k) partners |= (k != lst[i] && flipr[(size_t)i * nsm + k] >= 0);
    if (partners) {
      info[i] |= INFO_PARTNERS;
      plist.push_back(i);
    }
  }
  for (int i = 0; i <= nn; ++i)
    if (stab_ptr[i] < 0 || stab_ptr[i] > nstab || (i > 0 && stab_ptr[i] < stab_ptr[i - 1]))
      return fail("set_tables: stabiliser pointers");
  for (int j = 0; j < nstab; ++j)
    if (stab[j] < 0 || stab[j] >= op->nop) return fail("set_tables: stabiliser operation out of range");
  if (!op->hop && (dalloc(&op->hop, hop.size()) || dalloc(&op->flip, flip.size()) || dalloc(&op->info, nn) ||
                   dalloc(&op->sq, nn + 1) || dalloc(&op->isq, nn + 1) || dalloc(&op->stab_ptr, nn + 1)))
    return 1;
  if (op->stab) (void)hipFree(op->stab);
  op->stab = nullptr;
  if (dalloc(&op->stab, std::max(nstab, 1))) return 1;
  op->nstab = nstab;
  HCHECK(hipMemcpy(op->hop, hop.data(), hop.size() * 4, hipMemcpyHostToDevice));
  HCHECK(hipMemcpy(op->flip, flip.data(), flip.size() * 4, hipMemcpyHostToDevice));
  HCHECK(hipMemcpy(op->info, info.data(), info.size() * 4, hipMemcpyHostToDevice));
  if (op->plist) (void)hipFree(op->plist);
  op->plist = nullptr;
  op->nplist = (int)plist.size();
  if (dalloc(&op->plist, std::max(op->nplist, 1))) return 1;
  if (op->nplist) HCHECK(hipMemcpy(op->plist, plist.data(), plist.size() * 4, hipMemcpyHostToDevice));
  HCHECK(hipMemcpy(op->sq, sq, (size_t)nn * 8, hipMemcpyHostToDevice));
  {
    const double one = 1.0;
    HCHECK(hipMemcpy(op->sq + nn, &one, 8, hipMemcpyHostToDevice));
    std::vector<double> isq(nn + 1, 1.0);
    for (int i = 0; i < nn; ++i) {
      if (!(sq[i] > 0.0)) return fail("set_tables: sq must be positive");
      isq[i] = 1.0 / sq[i];
    }
    HCHECK(hipMemcpy(op->isq, isq.data(), isq.size() * 8, hipMemcpyHostToDevice));
  }
  HCHECK(hipMemcpy(op->stab_ptr, stab_ptr, (size_t)(nn + 1) * 4, hipMemcpyHostToDevice));
  if (nstab > 0) HCHECK(hipMemcpy(op->stab, stab, (size_t)nstab * 4, hipMemcpyHostToDevice));
  op->have_tables = true;
  return 0;
}

int asr_gpu_op_set_orbital_ops(asr_gpu_op* op, const int* pi, const double* sg) {
  if (!op) return fail("asr_gpu_op_set_orbital_ops: null operator");
  for (int q = 0; q < op->nop * 9; ++q)
    if (pi[q] < 0 || pi[q] > 8) return fail("set_orbital_ops: orbital index out of range");
  op->pi_h.assign(pi, pi + op->nop * 9);
  op->sg_h.assign(sg, sg + op->nop * 9);
  op->piid = true;
  for (int q = 0; q < op->nop * 9; ++q)
    if (pi[q] != q % 9) op->piid = false;
  std::vector<uint32_t> sgm(op->nop, 0u);
  for (int g = 0; g < op->nop; ++g)
    for (int l = 0; l < 9; ++l) {
      if (sg[g * 9 + l] != 1.0 && sg[g * 9 + l] != -1.0) return fail("set_orbital_ops: signs must be +-1 (monomial tables)");
      if (sg[g * 9 + l] < 0.0) sgm[g] |= 1u << l;
    }
  if (!op->pi && (dalloc(&op->pi, op->nop * 9) || dalloc(&op->sg, op->nop * 9) || dalloc(&op->sgm, op->nop))) return 1;
  HCHECK(hipMemcpy(op->pi, pi, (size_t)op->nop * 9 * 4, hipMemcpyHostToDevice));
  HCHECK(hipMemcpy(op->sg, sg, (size_t)op->nop * 9 * 8, hipMemcpyHostToDevice));
  HCHECK(hipMemcpy(op->sgm, sgm.data(), sgm.size() * 4, hipMemcpyHostToDevice));
  op->have_ops = true;
  return 0;
}

int asr_gpu_op_set_blocks(asr_gpu_op* op, const double* block, const double* block0) {
  if (!op) return fail("asr_gpu_op_set_blocks: null operator");
  const int ns = op->nslot;
  if ((size_t)op->ntype * (ns + 1) * 81 * 8 > 120 * 1024) return fail("set_blocks: bond tables exceed the LDS budget");
  std::vector<double> blk((size_t)op->ntype * (ns + 1) * 81);
  for (int t = 0; t < op->ntype; ++t) {
    std::memcpy(&blk[(size_t)t * (ns + 1) * 81], block0 + (size_t)t * 81, 81 * 8);
    std::memcpy(&blk[((size_t)t * (ns + 1) + 1) * 81], block + (size_t)t * ns * 81, (size_t)ns * 81 * 8);
  }
  // the padded rows, then (k_tile's scalar path, one site type) the same blocks transposed: column l2 at l2 * 10
  const size_t nb90 = (size_t)op->ntype * (ns + 1) * 90;
  std::vector<double> blkp(2 * nb90, 0.0);
  for (size_t b = 0; b < (size_t)op->ntype * (ns + 1); ++b)
    for (int l = 0; l < 9; ++l)
      for (int l2 = 0; l2 < 9; ++l2) {
        blkp[b * 90 + l * 10 + l2] = blk[b * 81 + l * 9 + l2];
        blkp[nb90 + b * 90 + l2 * 10 + l] = blk[b * 81 + l * 9 + l2];
      }
  if (!op->blk && (dalloc(&op->blk, blk.size()) || dalloc(&op->blkp, blkp.size()))) return 1;
  HCHECK(hipMemcpy(op->blk, blk.data(), blk.size() * 8, hipMemcpyHostToDevice));
  HCHECK(hipMemcpy(op->blkp, blkp.data(), blkp.size() * 8, hipMemcpyHostToDevice));
  op->have_blocks = true;
  return 0;
}

int asr_gpu_op_set_classes(asr_gpu_op* op, const int* nsc, const int* spc, const double* mt) {
  if (!op) return fail("asr_gpu_op_set_classes: null operator");
  for (int k = 0; k < op->ncls; ++k) {
    if (nsc[k] < 1 || nsc[k] > op->nsmax) return fail("set_classes: species count of a class out of range");
    for (int s = 0; s < nsc[k]; ++s)
      if (spc[k * op->nsmax + s] < 0 || spc[k * op->nsmax + s] >= op->nsp) return fail("set_classes: species out of range");
  }
  op->nsc_h.assign(nsc, nsc + op->ncls);
  op->spc_h.assign(spc, spc + (size_t)op->ncls * op->nsmax);
  op->mt_h.assign(mt, mt + (size_t)op->ncls * op->nsmax * op->nsmax * op->nsmax);
  if (!op->nsc && dalloc(&op->nsc, op->ncls)) return 1;
  HCHECK(hipMemcpy(op->nsc, nsc, (size_t)op->ncls * 4, hipMemcpyHostToDevice));
  op->have_classes = true;
  return 0;
}

// Member tile of the batched chains: 1, 2, ..., 64 members per wave (0 = autotuned per member count, the default).
int asr_gpu_op_set_tile(asr_gpu_op* op, int mt) {
  if (!op) return fail("asr_gpu_op_set_tile: null operator");
  if (mt != 65 && (mt < 0 || mt > 64 || (mt & (mt - 1)) != 0))
    return fail("asr_gpu_op_set_tile: 0, a power of two <= 64, or 65 (64 with complex members as lane pairs)");
  op->tile_pref = mt;
  return 0;
}

// The member tile the autotuner picked for nchain members (0 if no batched run of that size has happened).
// Target states of the projected chains (tnode / tgop of build_augcls_operator with tsite, src/asr_augcls.f90:805-819):
// ntarget x nsmax node indices (-1: none) and operations.
int asr_gpu_op_set_targets(asr_gpu_op* op, int ntarget, const int* tnode, const int* tgop) {
  if (!op || ntarget < 0 || (ntarget > 0 && (!tnode || !tgop))) return fail("asr_gpu_op_set_targets: arguments");
  const size_t n = (size_t)ntarget * op->nsmax;
  for (size_t q = 0; q < n; ++q)
    if (tnode[q] >= op->nn || (tnode[q] >= 0 && (tgop[q] < 0 || tgop[q] >= op->nop)))
      return fail("asr_gpu_op_set_targets: target node or operation out of range");
  if (op->tnode) (void)hipFree(op->tnode);
  if (op->tgop) (void)hipFree(op->tgop);
  op->tnode = op->tgop = nullptr;
  op->ntarget = 0;
  if (ntarget == 0) return 0;
  if (dalloc(&op->tnode, n) || dalloc(&op->tgop, n)) return 1;
  HCHECK(hipMemcpy(op->tnode, tnode, n * 4, hipMemcpyHostToDevice));
  HCHECK(hipMemcpy(op->tgop, tgop, n * 4, hipMemcpyHostToDevice));
  op->ntarget = ntarget;
  return 0;
}

// Projected chains of route edc (edc_green_proj, src/asr_edc.f90:174-232): the chains of asr_gpu_chains_edc run all nlev
// levels, every level's vector projected on the target states with the dressing weights tw (nchain x ntarget x nsmax x 9,
// spectral_dressing); on the host g = cf_species (the operator's tail) and gp = sum_k cpr_k pv_k, gp4 the same for the
// chain cut four levels earlier.  g: nchain complex; gp, gp4: nchain x ntarget x 9 complex.
int asr_gpu_chains_edc_proj(asr_gpu_op* op, int nchain, const int* start_orb, const double* params, const double* pab,
                            const int* frozen_sp, const double* conc, const double* z, const int* nlev, int project, int ldab,
                            const double* tw, double* a, double* b, int* mused, double* g, double* gp, double* gp4, int mode,
                            double* elapsed_ms) {
  if (!op) return fail("asr_gpu_chains_edc_proj: null operator");
  if (nchain <= 0) return 0;
  if (!op || !start_orb || !params || !pab || !z || !nlev || !tw || !a || !b || !mused || !g || !gp || !gp4)
    return fail("asr_gpu_chains_edc_proj: null argument");
  if (op->ntarget <= 0) return fail("asr_gpu_chains_edc_proj: no target states (asr_gpu_op_set_targets)");
  if (mode < 0 || mode > 2) return fail("asr_gpu_chains_edc_proj: mode 0 (batched), 1 (plain) or 2 (graph)");
  for (int j = 0; j < nchain; ++j)
    if (nlev[j] < 1 || nlev[j] > ldab) return fail("asr_gpu_chains_edc_proj: 1 <= nlev <= ldab");
  const int nt = op->ntarget;
  std::vector<cplx> cpr((size_t)nchain * ldab * nt * 9);
  double ms = 0.0;
  cplx* ac = (cplx*)a;
  cplx* bc = (cplx*)b;
  if (run_proj(op, nchain, start_orb, params, pab, frozen_sp, conc, z, nlev, tw, project, ldab, ac, bc, mused, cpr.data(),
               mode, &ms))
    return 1;
  if (elapsed_ms) *elapsed_ms = ms;
  std::vector<double> c0(2 * (size_t)ldab, 0.0), pv(2 * (size_t)ldab);
  for (int j = 0; j < nchain; ++j) {
    const int m = mused[j];
    const double* aj = (const double*)(ac + (size_t)j * ldab);
    const double* bj = (const double*)(bc + (size_t)j * ldab);
    for (int pass = 0; pass < 2; ++pass) {
      const int mm = pass == 0 ? m : std::max(1, m - 4);
      double gg[2], f[2];
      if (asr_host_cf_species(aj, bj, c0.data(), mm, z + 2 * j, op->depth_tail, gg, f, nullptr, pv.data()) < 0)
        return fail("asr_gpu_chains_edc_proj: continued fraction");
      if (pass == 0) { g[2 * j] = gg[0]; g[2 * j + 1] = gg[1]; }
      double* out = (pass == 0 ? gp : gp4) + (size_t)j * nt * 9 * 2;
      for (int e = 0; e < nt * 9; ++e) {
        cplx s{0.0, 0.0};
        for (int k = 0; k < mm; ++k) s = s + cpr[((size_t)j * ldab + k) * nt * 9 + e] * cplx{pv[2 * k], pv[2 * k + 1]};
        out[2 * e] = s.re;
        out[2 * e + 1] = s.im;
      }
    }
  }
  return 0;
}

// Full reorthogonalisation of the real chains (lanczos reorth = .true., src/reference/asr_recursion.f90:17-58).
int asr_gpu_op_set_compaction(asr_gpu_op* op, int on) {
  if (!op) return fail("asr_gpu_op_set_compaction: null handle");
  op->compact = on != 0;
  return 0;
}

int asr_gpu_op_set_window(asr_gpu_op* op, int nslot) {
  if (!op) return fail("asr_gpu_op_set_window: null handle");
  if (nslot < 0) return fail("asr_gpu_op_set_window: the number of slots must be >= 0");
  op->window = nslot;
  return 0;
}

int asr_gpu_op_set_window_queue(asr_gpu_op* op, int on_device) {
  if (!op) return fail("asr_gpu_op_set_window_queue: null handle");
  op->queue_dev = on_device != 0;
  return 0;
}

int asr_gpu_op_get_compactions(asr_gpu_op* op, int* n) {
  if (!op || !n) return fail("asr_gpu_op_get_compactions: null argument");
  *n = op->ncompact;
  return 0;
}

int asr_gpu_op_set_level_form(asr_gpu_op* op, int form) {
  if (!op) return fail("asr_gpu_op_set_level_form: null operator");
  if (form != 0 && form != 1) return fail("asr_gpu_op_set_level_form: form 0 (fused level) or 1 (separate update pass)");
  op->level_form = form;
  return 0;
}

int asr_gpu_op_set_reorth(asr_gpu_op* op, int reorth) {
  if (!op) return fail("asr_gpu_op_set_reorth: null operator");
  op->reorth = reorth != 0;
  return 0;
}

// Adaptive depth of the edc chains (ed_depth_tol / ed_tail of edc_green, src/asr_edc.f90:121-127): 0 = fixed depth.
int asr_gpu_op_set_depth_tol(asr_gpu_op* op, double depth_tol, int tail) {
  if (!op) return fail("asr_gpu_op_set_depth_tol: null operator");
  if (!(depth_tol >= 0.0) || tail < 0 || tail > 3) return fail("asr_gpu_op_set_depth_tol: depth_tol >= 0, tail 0..3");
  op->depth_tol = depth_tol;
  op->depth_tail = tail;
  return 0;
}

int asr_gpu_op_get_tile(asr_gpu_op* op, int nchain, int is_complex, int* mt) {
  if (!op) return fail("asr_gpu_op_get_tile: null operator");
  *mt = 0;
  for (auto& kv : op->tuned)
    if ((kv.first >> 8) == nchain && ((kv.first & 1) != 0) == (is_complex != 0)) *mt = kv.second;
  return 0;
}

int asr_gpu_op_trim(asr_gpu_op* op) {
  if (!op) return fail("asr_gpu_op_trim: null operator");
  if (op->stream) (void)hipStreamSynchronize(op->stream);
  op->pool.trim();
  return 0;
}

void asr_gpu_op_destroy(asr_gpu_op* op) {
  if (op && op->plist) (void)hipFree(op->plist);
  if (op && op->tnode) (void)hipFree(op->tnode);
  if (op && op->tgop) (void)hipFree(op->tgop);
  if (!op) return;
  for (void* p : {(void*)op->isq, (void*)op->sgm, (void*)op->blkp}) if (p) (void)hipFree(p);
  for (void* p : {(void*)op->hop, (void*)op->flip, (void*)op->info, (void*)op->sq, (void*)op->sg, (void*)op->blk,
                  (void*)op->stab_ptr, (void*)op->stab, (void*)op->pi, (void*)op->nsc})
    if (p) (void)hipFree(p);
  if (op->stream) (void)hipStreamDestroy(op->stream);
  delete op;
}

// y = H x for nb members (real vectors, host layout [member][9 nn]); the members' parameters as the chain calls.
int asr_gpu_apply_batched(asr_gpu_op* op, int nb, const int* start_orb, const double* params, const int* frozen_sp,
                          const double* conc, const double* x, double* y) {
  if (!op->have_tables || !op->have_ops || !op->have_blocks || !op->have_classes)
    return fail("asr_gpu_apply: operator tables incomplete");
  Members M;
  if (prepare_members(op, nb, start_orb, params, nullptr, frozen_sp, conc, nullptr, M)) return 1;
  MemBuffers MB;
  MemDev md;
  if (upload_members(op, M, MB, md)) return 1;
  const size_t dim = (size_t)op->nn * 9, nv = dim * M.nbp, nvp = nv + (size_t)9 * M.nbp;
  std::vector<double> hx;
  to_dev_layout(x, hx, nb, M.nbp, dim);
  double *dx, *dy, *t1, *t2, *h1;
  if (dalloc(&dx, nvp) || dalloc(&dy, nvp) || dalloc(&t1, nvp) || dalloc(&t2, nvp) || dalloc(&h1, nvp)) return 1;
  for (double* q : {dx, dy, t1, t2, h1}) HCHECK(hipMemsetAsync(q, 0, nvp * 8, op->stream));
  HCHECK(hipMemcpyAsync(dx, hx.data(), nv * 8, hipMemcpyHostToDevice, op->stream));
  int st = apply_chain_op<double>(op, M, md, dx, dy, t1, t2, h1, false, nullptr, nullptr, op->stream);
  if (!st) {
    HCHECK(hipMemcpyAsync(hx.data(), dy, nv * 8, hipMemcpyDeviceToHost, op->stream));
    HCHECK(hipStreamSynchronize(op->stream));
    from_dev_layout(hx, y, nb, M.nbp, dim);
  }
  for (double* p : {dx, dy, t1, t2, h1}) (void)hipFree(p);
  return st;
}

// w = X(z) x (augcls_rand_c with X(z) = z pa + pb) plus H x: the edc operator K(z) x for complex members.
int asr_gpu_apply_complex(asr_gpu_op* op, int nb, const int* start_orb, const double* params, const double* pab,
                          const int* frozen_sp, const double* conc, const double* z, const double* x, double* y) {
  if (!op->have_tables || !op->have_ops || !op->have_blocks || !op->have_classes)
    return fail("asr_gpu_apply: operator tables incomplete");
  Members M;
  if (prepare_members(op, nb, start_orb, params, pab, frozen_sp, conc, z, M)) return 1;
  MemBuffers MB;
  MemDev md;
  if (upload_members(op, M, MB, md)) return 1;
  const size_t dim = (size_t)op->nn * 9, nv = dim * M.nbp, nvp = nv + (size_t)9 * M.nbp;
  std::vector<cplx> hx;
  to_dev_layout((const cplx*)x, hx, nb, M.nbp, dim);
  cplx *dx, *dy, *t1, *t2, *h1;
  if (dalloc(&dx, nvp) || dalloc(&dy, nvp) || dalloc(&t1, nvp) || dalloc(&t2, nvp) || dalloc(&h1, nvp)) return 1;
  for (cplx* q : {dx, dy, t1, t2, h1}) HCHECK(hipMemsetAsync(q, 0, nvp * sizeof(cplx), op->stream));
  HCHECK(hipMemcpyAsync(dx, hx.data(), nv * sizeof(cplx), hipMemcpyHostToDevice, op->stream));
  int st = apply_chain_op<cplx>(op, M, md, dx, dy, t1, t2, h1, pab != nullptr, nullptr, nullptr, op->stream);
  if (!st) {
    HCHECK(hipMemcpyAsync(hx.data(), dy, nv * sizeof(cplx), hipMemcpyDeviceToHost, op->stream));
    HCHECK(hipStreamSynchronize(op->stream));
    from_dev_layout(hx, (cplx*)y, nb, M.nbp, dim);
  }
  for (cplx* p : {dx, dy, t1, t2, h1}) (void)hipFree(p);
  return st;
}

int asr_gpu_lanczos_real(asr_gpu_op* op, int nchain, const int* start_orb, const double* params, const int* frozen_sp,
                         const double* conc, const int* nlev, double tiny, int ldab, double* a, double* b, int* mused, int mode,
                         double* elapsed_ms) {
  if (!op) return fail("asr_gpu_lanczos_real: null operator");
  if (elapsed_ms) *elapsed_ms = 0.0;
  return run_chains<double>(op, nchain, start_orb, params, nullptr, frozen_sp, conc, nullptr, nlev, 0, ldab, a, b, mused, mode,
                            tiny > 0 ? tiny : 1e-12, true, elapsed_ms);
}

int asr_gpu_chains_edc(asr_gpu_op* op, int nchain, const int* start_orb, const double* params, const double* pab,
                       const int* frozen_sp, const double* conc, const double* z, const int* nlev, int project, int ldab,
                       double* a, double* b, int* mused, int mode, double* elapsed_ms) {
  if (!op) return fail("asr_gpu_chains_edc: null operator");
  if (elapsed_ms) *elapsed_ms = 0.0;
  return run_chains<cplx>(op, nchain, start_orb, params, pab, frozen_sp, conc, z, nlev, project, ldab, (cplx*)a, (cplx*)b,
                          mused, mode, 1e-12, false, elapsed_ms);
}

int asr_gpu_chains_edc_multi(int ngroup, asr_gpu_op* const* ops, const int* nchain, const int* start_orb,
                             const double* params, const double* pab, const int* frozen_sp, const double* conc, const double* z,
                             const int* nlev, int project, int ldab, double* a, double* b, int* mused, int mode,
                             double* elapsed_ms) {
  for (int g = 0; g < ngroup; ++g)
    if (!ops || !ops[g]) return fail("asr_gpu_chains_edc_multi: null operator");
  if (ngroup <= 0) return fail("chains_multi: no groups");
  return run_multi<cplx>(ngroup, ops, nchain, start_orb, params, pab, frozen_sp, conc, z, nlev, project, ldab, (cplx*)a,
                         (cplx*)b, mused, mode, 1e-12, false, elapsed_ms);
}

// The blocks of the two chains of every input block j (members 2j, 2j + 1): q0 = na^2 g_a, q1a = na f_a, q2 = nv^2 g_v,
// q1b = nv f_v with (g, f) from cf_species (the operator's tail) or, for a chain that closed, the finite fraction;
// bk = (q0 - 2 o_c q1 + o_c^2 q2, q1 - o_c q2, q2, q1a - q1b), q1 = (q1a + q1b) / 2 (ed3_blocks, src/asr_ed3.f90:247-278;
// with o_c = 0 and na = 1 the x of epi_blocks, src/asr_epi.f90:628-647).
static int finish_blocks(const asr_gpu_op* op, int nchain, const double* oc, const double* z, int ldab, const cplx* ac,
                         const cplx* bc, const cplx* ccv, const int* mused, const int* closed, const double* snorm,
                         double* bk) {
  for (int j = 0; j < nchain; ++j) {
    cplx q[2][2];  // (norm^2 g, norm f) per side
    for (int sd = 0; sd < 2; ++sd) {
      const int m = 2 * j + sd;
      const size_t o = (size_t)m * ldab;
      double g[2], f[2];
      int st;
      if (closed[m])
        st = asr_host_cf_closed((const double*)(ac + o), (const double*)(bc + o), (const double*)(ccv + o), mused[m], z + 2 * j,
                                g, f);
      else  // the status of cf_species counts bp fallbacks (hull values then), not errors
        st = asr_host_cf_species((const double*)(ac + o), (const double*)(bc + o), (const double*)(ccv + o), mused[m],
                                 z + 2 * j, op->depth_tail, g, f, nullptr, nullptr) < 0;
      if (st) return fail("chain blocks: continued fraction");
      const double nr = snorm[m];
      q[sd][0] = (nr * nr) * cplx{g[0], g[1]};
      q[sd][1] = nr * cplx{f[0], f[1]};
    }
    const cplx q0 = q[0][0], q1a = q[0][1], q2 = q[1][0], q1b = q[1][1];
    const double ocj = oc ? oc[j] : 0.0;
    const cplx fm = 0.5 * (q1a + q1b);
    const cplx r[4] = {q0 - (2.0 * ocj) * fm + (ocj * ocj) * q2, fm - ocj * q2, q2, q1a - q1b};
    for (int t = 0; t < 4; ++t) {
      bk[(size_t)j * 8 + 2 * t] = r[t].re;
      bk[(size_t)j * 8 + 2 * t + 1] = r[t].im;
    }
  }
  return 0;
}

// Route ed3 (ed3_blocks / ed3_blocks_blk, src/asr_ed3.f90:247-340): per input chain (start orbital, energy, central species)
// the chain from |a> / na projected on |v> and the chain from |v> / nv projected on |a>, run as one batch of 2 nchain
// members; the central blocks bk = (q0 - 2 o_c q1 + o_c^2 q2, q1 - o_c q2, q2, q1a - q1b) with q1 = (q1a + q1b) / 2 from
// cf_species (the operator's tail) or cf_closed.  params: c = g, d = D_c^1/2, o = e = 0 (ed3_setup); cf: nchain x 6 x nsp
// x 9 (et = z cf1 + cf2, u = z cf3 + cf4, Qt = z cf5 + cf6).
int asr_gpu_ed3_blocks(asr_gpu_op* op, int nchain, const int* start_orb, const double* params, const double* cf,
                       const double* oc, const int* frozen_sp, const double* conc, const double* z, const int* nlev,
                       int project, int ldab, double* bk, double* a, double* b, double* c, int* mused, int mode,
                       double* elapsed_ms) {
  if (!op) return fail("asr_gpu_ed3_blocks: null operator");
  if (nchain <= 0) return 0;
  if (!op || !start_orb || !params || !cf || !oc || !z || !nlev || !bk || !a || !b || !c || !mused)
    return fail("asr_gpu_ed3_blocks: null argument");
  if (mode < 0 || mode > 2) return fail("asr_gpu_ed3_blocks: mode 0 (batched), 1 (plain) or 2 (graph)");
  const int nsp9 = op->nsp * 9, ncn = op->ncls * op->nsmax, nm = 2 * nchain;
  for (int j = 0; j < nchain; ++j) {
    for (int q = 2 * nsp9; q < 4 * nsp9; ++q)
      if (params[(size_t)j * 4 * nsp9 + q] != 0.0) return fail("asr_gpu_ed3_blocks: o and e must be zero (ed3_setup)");
    if (nlev[j] < 1 || nlev[j] > ldab) return fail("asr_gpu_ed3_blocks: 1 <= nlev <= ldab");
  }
  std::vector<int> so(nm), fz(nm), nl(nm), side(nm);
  std::vector<double> pp((size_t)nm * 4 * nsp9), pab((size_t)nm * 2 * nsp9), pex((size_t)nm * 4 * nsp9), ocm(nm), zz(2 * nm),
      cc(conc ? (size_t)nm * ncn : 0);
  for (int j = 0; j < nchain; ++j)
    for (int sd = 0; sd < 2; ++sd) {
      const int q = 2 * j + sd;
      so[q] = start_orb[j];
      fz[q] = frozen_sp ? frozen_sp[j] : -1;
      nl[q] = nlev[j];
      side[q] = sd;
      ocm[q] = oc[j];
      zz[2 * q] = z[2 * j];
      zz[2 * q + 1] = z[2 * j + 1];
      std::memcpy(&pp[(size_t)q * 4 * nsp9], params + (size_t)j * 4 * nsp9, sizeof(double) * 4 * nsp9);
      if (conc) std::memcpy(&cc[(size_t)q * ncn], conc + (size_t)j * ncn, sizeof(double) * ncn);
      const double* cj = cf + (size_t)j * 6 * nsp9;
      // z - et = z (1 - cf1) - cf2; u = z cf3 + cf4; Qt = z cf5 + cf6
      for (int t = 0; t < nsp9; ++t) {
        pab[(size_t)q * 2 * nsp9 + t] = 1.0 - cj[t];
        pab[(size_t)q * 2 * nsp9 + nsp9 + t] = -cj[nsp9 + t];
      }
      std::memcpy(&pex[(size_t)q * 4 * nsp9], cj + 2 * nsp9, sizeof(double) * 4 * nsp9);
    }
  std::vector<int> closed(nm, 0);
  std::vector<double> snorm(nm), tnorm(nm);
  const BlockIn e3{pex.data(), ocm.data(), side.data(), nullptr, nullptr, nullptr};
  double ms = 0.0;
  cplx* ac = (cplx*)a;
  cplx* bc = (cplx*)b;
  cplx* ccv = (cplx*)c;
  if (run_blocks(op, nm, so.data(), pp.data(), pab.data(), fz.data(), conc ? cc.data() : nullptr, zz.data(), nl.data(), e3,
              project, ldab, ac, bc, ccv, mused, closed.data(), snorm.data(), tnorm.data(), mode, &ms))
    return 1;
  if (elapsed_ms) *elapsed_ms = ms;
  return finish_blocks(op, nchain, oc, z, ldab, ac, bc, ccv, mused, closed.data(), snorm.data(), bk);
}

// EPI pair blocks on the route-edc operator (epi_blocks / epi_blocks_blk, src/asr_epi.f90:628-647,740-759; the chains
// chain / chain_blk :590-626,650-737; epi_start :578-587): per block j the symmetry-adapted start vector v9 (unit norm,
// characters chi under the pair group) on the start node, the chain from it projected on |v> = h'|v9>, and the chain
// from |v> / nv projected on it; x = (X0, X1, X2, X1a - X1b) with X0 = <v9|G|v9>, X1 = <v9|h'G|v9>, X2 = <v9|h'Gh'|v9>.
// params, pab, z: the edc chain parameters of epi_setup (site 0 on the path point: its species and the partner's frozen
// through frozen_sp (fz slot 1) and frozen_sp2 (fz slot 2)).
int asr_gpu_epi_blocks(asr_gpu_op* op, int nchain, const double* v9, const double* chi, const double* params,
                       const double* pab, const int* frozen_sp, const int* frozen_sp2, const double* conc, const double* z,
                       const int* nlev, int project, int ldab, double* x, double* a, double* b, double* c, int* mused,
                       int mode, double* elapsed_ms) {
  if (!op) return fail("asr_gpu_epi_blocks: null operator");
  if (nchain <= 0) return 0;
  if (!op || !v9 || !chi || !params || !pab || !z || !nlev || !x || !a || !b || !c || !mused)
    return fail("asr_gpu_epi_blocks: null argument");
  if (mode < 0 || mode > 2) return fail("asr_gpu_epi_blocks: mode 0 (batched), 1 (plain) or 2 (graph)");
  const int nsp9 = op->nsp * 9, ncn = op->ncls * op->nsmax, nm = 2 * nchain;
  std::vector<int> so(nm, 0), fz(nm), fz2(nm), nl(nm), side(nm);
  std::vector<double> pp((size_t)nm * 4 * nsp9), pb((size_t)nm * 2 * nsp9), ocm(nm, 0.0), zz(2 * nm), vv((size_t)nm * 9),
      cc(conc ? (size_t)nm * ncn : 0);
  std::vector<uint64_t> chm(nm, 0);
  for (int j = 0; j < nchain; ++j) {
    if (nlev[j] < 1 || nlev[j] > ldab) return fail("asr_gpu_epi_blocks: 1 <= nlev <= ldab");
    uint64_t mk = 0;
    for (int g = 0; g < op->nop; ++g) {
      const double ch = chi[(size_t)j * op->nop + g];
      if (std::fabs(std::fabs(ch) - 1.0) > 1e-12) return fail("augcls_set_chi: characters must be +-1");
      if (ch < 0.0) mk |= 1ull << g;
    }
    if (chi[(size_t)j * op->nop] < 0.0) return fail("augcls_set_chi: chi(identity) must be 1");
    for (int sd = 0; sd < 2; ++sd) {
      const int q = 2 * j + sd;
      fz[q] = frozen_sp ? frozen_sp[j] : -1;
      fz2[q] = frozen_sp2 ? frozen_sp2[j] : -1;
      nl[q] = nlev[j];
      side[q] = sd;
      chm[q] = mk;
      zz[2 * q] = z[2 * j];
      zz[2 * q + 1] = z[2 * j + 1];
      std::memcpy(&vv[(size_t)q * 9], v9 + (size_t)j * 9, sizeof(double) * 9);
      std::memcpy(&pp[(size_t)q * 4 * nsp9], params + (size_t)j * 4 * nsp9, sizeof(double) * 4 * nsp9);
      std::memcpy(&pb[(size_t)q * 2 * nsp9], pab + (size_t)j * 2 * nsp9, sizeof(double) * 2 * nsp9);
      if (conc) std::memcpy(&cc[(size_t)q * ncn], conc + (size_t)j * ncn, sizeof(double) * ncn);
    }
  }
  std::vector<int> closed(nm, 0);
  std::vector<double> snorm(nm), tnorm(nm);
  const BlockIn e3{nullptr, ocm.data(), side.data(), vv.data(), chm.data(), fz2.data()};
  double ms = 0.0;
  cplx* ac = (cplx*)a;
  cplx* bc = (cplx*)b;
  cplx* ccv = (cplx*)c;
  if (run_blocks(op, nm, so.data(), pp.data(), pb.data(), fz.data(), conc ? cc.data() : nullptr, zz.data(), nl.data(), e3,
                 project, ldab, ac, bc, ccv, mused, closed.data(), snorm.data(), tnorm.data(), mode, &ms))
    return 1;
  if (elapsed_ms) *elapsed_ms = ms;
  return finish_blocks(op, nchain, nullptr, z, ldab, ac, bc, ccv, mused, closed.data(), snorm.data(), x);
}

int asr_gpu_lanczos_real_multi(int ngroup, asr_gpu_op* const* ops, const int* nchain, const int* start_orb,
                               const double* params, const int* frozen_sp, const double* conc, const int* nlev, double tiny,
                               int ldab, double* a, double* b, int* mused, int mode, double* elapsed_ms) {
  for (int g = 0; g < ngroup; ++g)
    if (!ops || !ops[g]) return fail("asr_gpu_lanczos_real_multi: null operator");
  if (ngroup <= 0) return fail("lanczos_multi: no groups");
  return run_multi<double>(ngroup, ops, nchain, start_orb, params, nullptr, frozen_sp, conc, nullptr, nlev, 0, ldab, a, b,
                           mused, mode, tiny > 0 ? tiny : 1e-12, true, elapsed_ms);
}

int asr_gpu_lanczos_adjoint(asr_gpu_op* op, int nchain, const int* start_orb, const double* params, const int* frozen_sp,
                            const double* conc, int nlev, int ckpt, const double* dl_da, const double* dl_db, double* a,
                            double* b, double* grad, double* fwd_ms, double* adj_ms) {
  if (!op) return fail("asr_gpu_lanczos_adjoint: null operator");
  if (nchain <= 0 || nlev <= 0) return fail("lanczos_adjoint: no chains");
  return run_adjoint(op, nchain, start_orb, params, frozen_sp, conc, nlev, ckpt, dl_da, dl_db, a, b, grad, fwd_ms, adj_ms);
}

// Per-kernel durations of the chain level step: the same chains as asr_gpu_chains_edc (pab != NULL) or
// asr_gpu_lanczos_real (pab == NULL), every kernel of every level bracketed by HIP events on the engine's stream.
// kernel_ms[q], kind[q] (q < *nk <= 16): mean duration of the q-th launch of a level and its kind
// (0 fused SpMM, 1 flip-gather term, 2 a reduction, 3 Lanczos update + projection, 4 b reduction, 5 scaling).
int asr_gpu_profile_chains(asr_gpu_op* op, int nchain, const int* start_orb, const double* params, const double* pab,
                           const int* frozen_sp, const double* conc, const double* z, const int* nlev, int project, int* nk,
                           double* kernel_ms, int* kind) {
  if (!op) return fail("asr_gpu_profile_chains: null operator");
  KTimer T;
  g_timer = &T;
  g_acc.clear();
  g_cnt.clear();
  g_levels = 0;
  int maxd = 0;
  for (int c = 0; c < nchain; ++c) maxd = std::max(maxd, nlev[c]);
  std::vector<double> a((size_t)nchain * maxd * 2), b((size_t)nchain * maxd * 2);
  std::vector<int> mu(nchain);
  double ms = 0.0;
  int st;
  if (pab)
    st = run_chains<cplx>(op, nchain, start_orb, params, pab, frozen_sp, conc, z, nlev, project, maxd, (cplx*)a.data(),
                          (cplx*)b.data(), mu.data(), 0, 1e-12, false, &ms);
  else
    st = run_chains<double>(op, nchain, start_orb, params, nullptr, frozen_sp, conc, nullptr, nlev, 0, maxd, a.data(),
                            b.data(), mu.data(), 0, 1e-12, true, &ms);
  g_timer = nullptr;
  for (hipEvent_t e : T.ev) (void)hipEventDestroy(e);
  if (st) return st;
  int n = std::min((int)g_acc.size(), 16);
  *nk = n;
  for (int q = 0; q < n; ++q) {
    kernel_ms[q] = g_cnt[q] ? g_acc[q] / g_cnt[q] : 0.0;
    kind[q] = T.kind[q];
  }
  return 0;
}

// ------------------------------------------------------------------------------------------------ KPM
// the per-species tables of the Chebyshev steps, 1 / d and (c - cshift) / d^2 for the shift cshift, copied to the device
// (ordered after the handle's earlier work, the host copy read before the call returns)
static int real_shift(asr_gpu_real* R, double cshift) {
  const int n = R->nsp * 9;
  for (int q = 0; q < n; ++q) {
    const double c = R->par_h[q], d = R->par_h[n + q];
    R->tab_h[q] = 1.0 / d;
    R->tab_h[n + q] = (c - cshift) / (d * d);
  }
  HCHECK(hipMemcpyAsync(R->tab, R->tab_h.data(), (size_t)2 * n * 8, hipMemcpyHostToDevice, R->stream));
  HCHECK(hipStreamSynchronize(R->stream));
  return 0;
}

int asr_gpu_real_create(int nsite, int nslot, const int* nbr, int nsp, const signed char* species, const double* par,
                        const double* block, const double* block0, asr_gpu_real** out) {
  *out = nullptr;
  if (nslot > 18) return fail("asr_gpu_real_create: at most 18 neighbour slots");
  for (size_t q = 0; q < (size_t)nsite * nslot; ++q)
    if (nbr[q] < 0 || nbr[q] >= nsite) return fail("asr_gpu_real_create: neighbour index out of range");
  for (int i = 0; i < nsite; ++i)
    if (species[i] < 0 || species[i] >= nsp) return fail("asr_gpu_real_create: species out of range");
  for (int q = 2 * nsp * 9; q < 4 * nsp * 9; ++q)
    if (par[q] != 0.0) return fail("asr_gpu_real_create: the KPM kernel applies h = c + d S d (o = E_nu = 0)");
  auto* R = new asr_gpu_real();
  R->nsite = nsite; R->nslot = nslot; R->nsp = nsp;
  std::vector<double> blk((size_t)(nslot + 1) * 81);
  std::memcpy(blk.data(), block0, 81 * 8);
  std::memcpy(blk.data() + 81, block, (size_t)nslot * 81 * 8);
  if (dalloc(&R->nbr, (size_t)nsite * nslot) || dalloc(&R->spec, nsite) || dalloc(&R->par, (size_t)4 * nsp * 9) ||
      dalloc(&R->blk, blk.size()) || dalloc(&R->tab, (size_t)2 * nsp * 9)) {
    delete R;
    return 1;
  }
  HCHECK(hipStreamCreateWithFlags(&R->stream, hipStreamNonBlocking));
  HCHECK(hipGetDevice(&R->device));
  double* sp = nullptr;  // the device address of the no-alloc entry's scratch, looked up at set-up
  if (scratch_ptr(R->device, &sp)) return 1;
  HCHECK(hipMemcpy(R->nbr, nbr, (size_t)nsite * nslot * 4, hipMemcpyHostToDevice));
  HCHECK(hipMemcpy(R->spec, species, nsite, hipMemcpyHostToDevice));
  HCHECK(hipMemcpy(R->par, par, (size_t)4 * nsp * 9 * 8, hipMemcpyHostToDevice));
  HCHECK(hipMemcpy(R->blk, blk.data(), blk.size() * 8, hipMemcpyHostToDevice));
  R->par_h.assign(par, par + (size_t)4 * nsp * 9);
  R->tab_h.assign((size_t)2 * nsp * 9, 0.0);
  for (int q = 0; q < nsp * 9; ++q)
    if (!(par[nsp * 9 + q] != 0.0)) return fail("asr_gpu_real_create: d must be non-zero");
  if (real_shift(R, 0.0)) return 1;
  *out = R;
  return 0;
}

void asr_gpu_real_destroy(asr_gpu_real* R) {
  if (!R) return;
  for (void* p : {(void*)R->nbr, (void*)R->spec, (void*)R->par, (void*)R->blk, (void*)R->tab})
    if (p) (void)hipFree(p);
  if (R->stream) (void)hipStreamDestroy(R->stream);
  delete R;
}

static RealDev real_dev(const asr_gpu_real* R) {
  RealDev d;
  d.nsite = R->nsite; d.nslot = R->nslot; d.nsp = R->nsp; d.nbr = R->nbr; d.spec = R->spec; d.par = R->par; d.blk = R->blk;
  d.idv = R->tab;
  d.fsh = R->tab + (size_t)R->nsp * 9;
  return d;
}

static int kpm_vt(const asr_gpu_real* R, int nvec) {
  int vt = R->vt > 0 ? R->vt : 64;
  while (vt > 8 && vt / 2 >= nvec) vt /= 2;
  return vt;
}

static int launch_kpm_step(const asr_gpu_real* R, int vt, dim3 grid, int nvec, int nvp, const double* x, const double* t0,
                           double* y, double c, double a, double b, uint64_t seed, double* part, int want_dot) {
  const RealDev rd = real_dev(R);
  switch (vt) {
    case 8: hipLaunchKernelGGL(k_kpm_step<8>, grid, dim3(BLOCK), 0, R->stream, rd, nvec, nvp, x, t0, y, c, a, b, seed, part, want_dot); break;
    case 16: hipLaunchKernelGGL(k_kpm_step<16>, grid, dim3(BLOCK), 0, R->stream, rd, nvec, nvp, x, t0, y, c, a, b, seed, part, want_dot); break;
    case 32: hipLaunchKernelGGL(k_kpm_step<32>, grid, dim3(BLOCK), 0, R->stream, rd, nvec, nvp, x, t0, y, c, a, b, seed, part, want_dot); break;
    default: hipLaunchKernelGGL(k_kpm_step<64>, grid, dim3(BLOCK), 0, R->stream, rd, nvec, nvp, x, t0, y, c, a, b, seed, part, want_dot); break;
  }
  HCHECK(hipGetLastError());
  return 0;
}

static dim3 kpm_grid(const asr_gpu_real* R, int vt, int nvp) {
  const int sites_per_block = (BLOCK / WAVE) * (WAVE / vt);
  return dim3((R->nsite + sites_per_block - 1) / sites_per_block, nvp / vt);
}

// kpm_moments by moment doubling (two moments per Chebyshev step)
int asr_gpu_real_set_doubling(asr_gpu_real* R, int on) {
  if (!R) return fail("asr_gpu_real_set_doubling: null handle");
  R->dbl = on != 0;
  return 0;
}

// Vector tile of the KPM step (vectors per site in a wave: 8, 16, 32 or 64; 0 = automatic).
int asr_gpu_real_set_tile(asr_gpu_real* R, int vt) {
  if (!R) return fail("asr_gpu_real_set_tile: null handle");
  if (vt < 0 || vt > 4)
    return fail("asr_gpu_real_set_tile: 0 (automatic), 1 or 2 vectors per lane, 3 = 16-vector slices, 4 = MFMA contraction");
  R->v2 = (vt != 1);
  R->s16 = (vt == 3);
  R->mf = (vt == 4);
  return 0;
}

// y = h x for nvec vectors (host layout [vector][9 nsite])
int asr_gpu_real_apply(asr_gpu_real* R, int nvec, const double* x, double* y) {
  if (!R) return fail("asr_gpu_real_apply: null handle");
  const size_t dim = (size_t)R->nsite * 9;
  const int vt = kpm_vt(R, nvec);
  const int nvp = ((nvec + vt - 1) / vt) * vt;
  std::vector<double> hx;
  to_dev_layout(x, hx, nvec, nvp, dim);
  double *dx, *dy;
  if (dalloc(&dx, dim * nvp) || dalloc(&dy, dim * nvp)) return 1;
  HCHECK(hipMemcpy(dx, hx.data(), dim * nvp * 8, hipMemcpyHostToDevice));
  if (launch_kpm_step(R, vt, kpm_grid(R, vt, nvp), nvec, nvp, dx, nullptr, dy, 0.0, 1.0, 0.0, 0ull, nullptr, 0)) return 1;
  HCHECK(hipStreamSynchronize(R->stream));
  HCHECK(hipMemcpy(hx.data(), dy, dim * nvp * 8, hipMemcpyDeviceToHost));
  from_dev_layout(hx, y, nvec, nvp, dim);
  (void)hipFree(dx);
  (void)hipFree(dy);
  return 0;
}

// The device buffers of one KPM recurrence: three Chebyshev vectors (dim x nvp), the per-block dot partials (and those of
// moment doubling), the per-vector moments (nvec x nmom).
struct KpmBufs {
  double *t0 = nullptr, *t1 = nullptr, *t2 = nullptr, *part = nullptr, *part2 = nullptr, *mper = nullptr, *stage = nullptr;
};
struct KpmSizes {
  size_t vec, part, part2, mper, stage;
  size_t total() const { return 3 * vec + part + part2 + mper + stage; }
};
static KpmSizes kpm_sizes(const asr_gpu_real* R, int nmom, int nvec) {
  const size_t dim = (size_t)R->nsite * 9;
  const int vt = 64;
  const int nvp = ((nvec + vt - 1) / vt) * vt;
  const dim3 grid = kpm_grid(R, vt, nvp);
  KpmSizes z;
  z.vec = dim * nvp;
  z.part = (size_t)grid.x * grid.y * vt;
  z.part2 = R->dbl ? (size_t)((R->nsite + 3) / 4) * grid.y * 64 : 0;
  z.mper = (size_t)nvec * nmom;
  z.stage = (size_t)2 * grid.y * KPS * vt;
  return z;
}
static int kpm_check(int nmom, int nvec, uint64_t seed, const asr_gpu_real* R) {
  if (nmom < 1 || nvec < 1) return fail("kpm_moments: nmom and nvec must be positive");
  if (seed >= (1ull << 24) || nvec > 4096) return fail("kpm_moments: seed < 2^24 and nvec <= 4096");
  if ((size_t)R->nsite * 9 >= (1ull << 28)) return fail("kpm_moments: more than 2^28 rows");
  return 0;
}

// The Chebyshev recurrence of kpm_moments on the handle's stream into the given buffers: mper[v][n] = <r_v|t_n> / dim.
// ev (2 + 2 x steps events, or null): the whole recurrence and each timed Chebyshev step.
static int kpm_run(asr_gpu_real* R, double c, double s, int nmom, int nvec, uint64_t seed, KpmBufs B, hipEvent_t* ev) {
  double *t0 = B.t0, *t1 = B.t1, *t2 = B.t2, *part = B.part, *part2 = B.part2, *mper = B.mper, *stage = B.stage;
  const size_t dim = (size_t)R->nsite * 9;
  const int vt = 64;
  const int nvp = ((nvec + vt - 1) / vt) * vt;
  const dim3 grid = kpm_grid(R, vt, nvp);
  const double inv_dim = 1.0 / (double)dim;
  const RealDev rd = real_dev(R);
  auto mark = [&](int q) -> int {
    if (ev) HCHECK(hipEventRecord(ev[q], R->stream));
    return 0;
  };
  // the moments of one step from its partials (nbx rows per vector tile; m < 0: mu_n, else the doubling pair of t_m)
  auto fin = [&](const double* p1, const double* p2, int nbx, int n, int m) {
    hipLaunchKernelGGL(k_kpm_psum, dim3(KPS, grid.y), dim3(WAVE * KPW), 0, R->stream, p1, p2, nbx, vt, stage);
    hipLaunchKernelGGL(k_kpm_pfin, dim3(grid.y), dim3(WAVE * KPF_WAVES), 0, R->stream, (const double*)stage, p2 ? 1 : 0, vt,
                       (int)grid.y,
                       nvec, inv_dim, mper, nmom, n, m);
  };
  if (real_shift(R, c)) return 1;
  if (mark(0)) return 1;
  // t0 = r (stored as d o r), mu_0 = <r|r>/dim: the step with alpha = 0, beta = -1
  hipLaunchKernelGGL((k_kpm_u<true, 0>), grid, dim3(BLOCK), 0, R->stream, rd, (const double*)R->blk, nvec, nvp,
                     (const double*)nullptr, (const double*)nullptr, t0, 0.0, 0.0, -1.0, seed, part, 1);
  fin(part, nullptr, (int)grid.x, 0, -1);
  if (nmom > 1) {
    // t1 = (H r - c r) / s
    hipLaunchKernelGGL((k_kpm_u<true, 0>), grid, dim3(BLOCK), 0, R->stream, rd, (const double*)R->blk, nvec, nvp,
                       (const double*)nullptr, (const double*)nullptr, t1, c, 1.0 / s, 0.0, seed, part, 1);
    fin(part, nullptr, (int)grid.x, 1, -1);
  }
  if (R->dbl) {
    // moment doubling: the step that makes t_m gives mu_2m and mu_2m-1, so t_2 .. t_M with 2 M >= nmom - 1 suffice
    const int nbx2 = (int)((R->nsite + 7) / 8);
    const dim3 g2((unsigned)nbx2, grid.y);
    if (nmom > 2) {
      hipLaunchKernelGGL(k_kpm_sq, g2, dim3(BLOCK), 0, R->stream, rd, nvp, (const double*)t1, part);
      fin(part, nullptr, nbx2, 0, 1);
    }
    const int M = nmom / 2;  // the largest m with 2 m - 1 <= nmom - 1
    const bool mfd = R->mf && R->nslot == 18;
    const int nbxm = mfd ? (int)((R->nsite + 3) / 4) : nbx2;
    for (int m = 2; m <= M; ++m) {
      if (mark(2 + 2 * (m - 2))) return 1;
      if (mfd)
        hipLaunchKernelGGL((k_kpm_mf<18, true>), dim3((unsigned)nbxm, grid.y), dim3(BLOCK), 0, R->stream, rd,
                           (const double*)R->blk, nvec, nvp, (const double*)t1, (const double*)t0, t2, c, 2.0 / s, 1.0, seed,
                           part, part2);
      else if (R->nslot == 18)
        hipLaunchKernelGGL((k_kpm_v2<18, true>), g2, dim3(BLOCK), 0, R->stream, rd, (const double*)R->blk, nvec, nvp,
                           (const double*)t1, (const double*)t0, t2, c, 2.0 / s, 1.0, seed, part, part2);
      else
        hipLaunchKernelGGL((k_kpm_v2<0, true>), g2, dim3(BLOCK), 0, R->stream, rd, (const double*)R->blk, nvec, nvp,
                           (const double*)t1, (const double*)t0, t2, c, 2.0 / s, 1.0, seed, part, part2);
      if (mark(3 + 2 * (m - 2))) return 1;
      fin(part, part2, nbxm, 0, m);
      double* tmp = t0; t0 = t1; t1 = t2; t2 = tmp;
    }
  }
  for (int n = 2; n < nmom && !R->dbl; ++n) {
    // t2 = 2 (H t1 - c t1) / s - t0
    if (mark(2 + 2 * (n - 2))) return 1;
    if (R->mf && R->nslot == 18)
      hipLaunchKernelGGL((k_kpm_mf<18, false>), dim3((R->nsite + 3) / 4, grid.y), dim3(BLOCK), 0, R->stream, rd,
                         (const double*)R->blk, nvec, nvp, (const double*)t1, (const double*)t0, t2, c, 2.0 / s, 1.0, seed,
                         part, (double*)nullptr);
    else if (R->s16 && R->nslot == 18)
      hipLaunchKernelGGL((k_kpm_s16<18>), dim3((R->nsite + 31) / 32, nvp / 16), dim3(BLOCK), (size_t)19 * 90 * 8, R->stream, rd,
                         nvec, nvp, (const double*)t1, (const double*)t0, t2, c, 2.0 / s, 1.0, seed, part);
    else if (R->v2 && R->nslot == 18)
      hipLaunchKernelGGL((k_kpm_v2<18, false>), dim3((R->nsite + 7) / 8, grid.y), dim3(BLOCK), 0, R->stream, rd,
                         (const double*)R->blk, nvec, nvp, (const double*)t1, (const double*)t0, t2, c, 2.0 / s, 1.0, seed,
                         part, (double*)nullptr);
    else if (R->v2)
      hipLaunchKernelGGL((k_kpm_v2<0, false>), dim3((R->nsite + 7) / 8, grid.y), dim3(BLOCK), 0, R->stream, rd,
                         (const double*)R->blk, nvec, nvp, (const double*)t1, (const double*)t0, t2, c, 2.0 / s, 1.0, seed,
                         part, (double*)nullptr);
    else if (R->nslot == 18)
      hipLaunchKernelGGL((k_kpm_u<false, 18>), grid, dim3(BLOCK), 0, R->stream, rd, (const double*)R->blk, nvec, nvp,
                         (const double*)t1, (const double*)t0, t2, c, 2.0 / s, 1.0, seed, part, 1);
    else
      hipLaunchKernelGGL((k_kpm_u<false, 0>), grid, dim3(BLOCK), 0, R->stream, rd, (const double*)R->blk, nvec, nvp,
                         (const double*)t1, (const double*)t0, t2, c, 2.0 / s, 1.0, seed, part, 1);
    if (mark(3 + 2 * (n - 2))) return 1;
    const int nbx = (R->mf && R->nslot == 18) ? (int)((R->nsite + 3) / 4)
                    : (R->s16 && R->nslot == 18) ? (int)((R->nsite + 31) / 32)
                    : R->v2 ? (int)((R->nsite + 7) / 8) : (int)grid.x;
    fin(part, nullptr, nbx, n, -1);
    double* tmp = t0; t0 = t1; t1 = t2; t2 = tmp;
  }
  if (mark(1)) return 1;
  HCHECK(hipGetLastError());
  return 0;
}

// mu_n = (1/R) sum_r <r|T_n((H - c)/s)|r> / dim (kpm_moments, src/reference/asr_kpm.f90:39-65) for nvec Rademacher vectors
// in one batched SpMM per moment.  mu_per (optional): nvec x nmom per-vector terms.  step_ms: mean duration of the
// Chebyshev step kernel alone (HIP events around each launch on the engine's stream); total_ms: the whole recurrence.
int asr_gpu_kpm_moments(asr_gpu_real* R, double c, double s, int nmom, int nvec, uint64_t seed, double* mu, double* mu_per,
                        double* step_ms, double* total_ms) {
  if (!R) return fail("asr_gpu_kpm_moments: null handle");
  if (kpm_check(nmom, nvec, seed, R)) return 1;
  const KpmSizes z = kpm_sizes(R, nmom, nvec);
  KpmBufs B;
  if (dalloc(&B.t0, z.vec) || dalloc(&B.t1, z.vec) || dalloc(&B.t2, z.vec) || dalloc(&B.part, z.part) ||
      dalloc(&B.mper, z.mper) || dalloc(&B.stage, z.stage))
    return 1;
  if (R->dbl && dalloc(&B.part2, z.part2)) return 1;
  // Chebyshev steps timed: t_2 .. t_(nmom-1), or t_2 .. t_(nmom/2) with the doubling
  const int nstep_ev = R->dbl ? std::max(nmom / 2 - 1, 0) : (nmom > 2 ? nmom - 2 : 0);
  std::vector<hipEvent_t> ev(2 + 2 * nstep_ev);
  for (auto& e : ev) HCHECK(hipEventCreate(&e));
  if (kpm_run(R, c, s, nmom, nvec, seed, B, ev.data())) return 1;
  HCHECK(hipEventSynchronize(ev[1]));
  float ms_all = 0.f;
  HCHECK(hipEventElapsedTime(&ms_all, ev[0], ev[1]));
  double ms_steps = 0.0;
  for (int q = 0; q < nstep_ev; ++q) {
    float t = 0.f;
    HCHECK(hipEventElapsedTime(&t, ev[2 + 2 * q], ev[3 + 2 * q]));
    ms_steps += t;
  }
  std::vector<double> per((size_t)nvec * nmom);
  HCHECK(hipMemcpy(per.data(), B.mper, per.size() * 8, hipMemcpyDeviceToHost));
  for (int n = 0; n < nmom; ++n) {
    double acc = 0.0;
    for (int v = 0; v < nvec; ++v) acc += per[(size_t)v * nmom + n];
    mu[n] = acc / (double)nvec;
  }
  if (mu_per) std::memcpy(mu_per, per.data(), per.size() * 8);
  if (step_ms) *step_ms = nstep_ev ? ms_steps / nstep_ev : 0.0;
  if (total_ms) *total_ms = ms_all;
  for (double* p : {B.t0, B.t1, B.t2, B.part, B.mper, B.stage}) (void)hipFree(p);
  if (B.part2) (void)hipFree(B.part2);
  for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  return 0;
}

int asr_gpu_kpm_moments_noalloc(asr_gpu_real* R, double c, double s, int nmom, int nvec, uint64_t seed, double* mu_dev) {
  if (!R || !mu_dev) return fail("asr_gpu_kpm_moments_noalloc: null handle or output");
  if (kpm_check(nmom, nvec, seed, R)) return 1;
  const KpmSizes z = kpm_sizes(R, nmom, nvec);
  if (z.total() > ASR_NOALLOC_SCRATCH_DOUBLES)
    return fail("asr_gpu_kpm_moments_noalloc: the work arrays (" + std::to_string(z.total() * 8) +
                " bytes: three Chebyshev vectors of 9 x sites x vectors rounded up to 64, partials, moments) exceed the "
                "fixed device scratch of " + std::to_string((size_t)ASR_NOALLOC_SCRATCH_DOUBLES * 8) + " bytes");
  double* base = nullptr;
  if (scratch_ptr(R->device, &base)) return 1;
  KpmBufs B;
  B.t0 = base;
  B.t1 = B.t0 + z.vec;
  B.t2 = B.t1 + z.vec;
  B.part = B.t2 + z.vec;
  B.part2 = R->dbl ? B.part + z.part : nullptr;
  B.mper = B.part + z.part + z.part2;
  B.stage = B.mper + z.mper;
  if (kpm_run(R, c, s, nmom, nvec, seed, B, nullptr)) return 1;
  hipLaunchKernelGGL(k_kpm_mean, dim3((nmom + 255) / 256), dim3(256), 0, R->stream, (const double*)B.mper, nvec, nmom, mu_dev);
  HCHECK(hipGetLastError());
  HCHECK(hipStreamSynchronize(R->stream));
  return 0;
}

// Extremal eigenvalues of h (spectral_bounds, src/reference/asr_kpm.f90:14-36): nstep Lanczos steps from nvec random
// start vectors at once (column v starts from splitmix64(((seed + v) << 40) ^ row), uniform in [-1/2, 1/2)); the
// eigenvalues of each chain's tridiagonal matrix (the role of dsterf) give emin = min - margin, emax = max + margin over
// the chains.  a, b: nvec x nstep or NULL (b(i) = |w_i| of every step, as the reference keeps it).
int asr_gpu_spectral_bounds(asr_gpu_real* R, int nstep, int nvec, uint64_t seed, double margin, double* emin, double* emax,
                            double* a_out, double* b_out) {
  if (!R || nstep < 1 || nvec < 1 || nvec > 64 || !emin || !emax) return fail("asr_gpu_spectral_bounds: arguments");
  const size_t nrow = (size_t)R->nsite * 9;
  int vt = 8;
  while (vt < nvec) vt <<= 1;
  const int nvp = vt;
  const size_t n = nrow * nvp;
  double *u = nullptr, *up = nullptr, *w = nullptr, *part = nullptr, *coef = nullptr;
  const int nbd = 1024;
  if (dalloc(&u, n) || dalloc(&up, n) || dalloc(&w, n) || dalloc(&part, (size_t)nbd * nvp) ||
      dalloc(&coef, 2 * nvp))
    return 1;
  auto cleanup = [&]() {
    for (double* p : {u, up, w, part, coef}) (void)hipFree(p);
  };
  std::vector<double> ph((size_t)nbd * nvp), av((size_t)nvec * nstep), bv((size_t)nvec * nstep), cf(2 * nvp, 0.0);
  const dim3 eg((unsigned)((n + 255) / 256));
  // per-column sums of x .* y on the host, fixed order
  auto coldot = [&](const double* x, const double* y, std::vector<double>& out) -> int {
    hipLaunchKernelGGL(k_col_dot, dim3(nbd), dim3(BLOCK), 0, R->stream, x, y, nrow, nvp, part);
    if (hipMemcpyAsync(ph.data(), part, ph.size() * 8, hipMemcpyDeviceToHost, R->stream) != hipSuccess ||
        hipStreamSynchronize(R->stream) != hipSuccess)
      return fail("spectral_bounds: reduction");
    out.assign(nvp, 0.0);
    for (int bq = 0; bq < nbd; ++bq)
      for (int v = 0; v < nvp; ++v) out[v] += ph[(size_t)bq * nvp + v];
    return 0;
  };
  auto upload = [&](const std::vector<double>& h) -> int {
    return hipMemcpyAsync(coef, h.data(), h.size() * 8, hipMemcpyHostToDevice, R->stream) != hipSuccess;
  };
  std::vector<double> dd;
  int st = 0;
  hipLaunchKernelGGL(k_rl_start, eg, dim3(256), 0, R->stream, u, nrow, nvp, nvec, seed);
  st |= hipMemsetAsync(up, 0, n * 8, R->stream) != hipSuccess;
  st |= hipMemsetAsync(w, 0, n * 8, R->stream) != hipSuccess;  // the step writes the columns v < nvec only
  if (!st) st = coldot(u, u, dd);
  if (!st) {
    for (int v = 0; v < nvp; ++v) cf[v] = v < nvec ? 1.0 / std::sqrt(dd[v]) : 0.0;
    st = upload(cf);
  }
  if (!st) hipLaunchKernelGGL(k_rl_shift, eg, dim3(256), 0, R->stream, u, up, (const double*)nullptr, (const double*)coef, n, nvp);
  std::vector<double> bn(nvp, 0.0);
  for (int i = 0; i < nstep && !st; ++i) {
    st = launch_kpm_step(R, vt, kpm_grid(R, vt, nvp), nvec, nvp, u, nullptr, w, 0.0, 1.0, 0.0, 0ull, nullptr, 0);
    if (!st) st = coldot(u, w, dd);
    if (st) break;
    for (int v = 0; v < nvp; ++v) {
      cf[v] = v < nvec ? dd[v] : 0.0;
      cf[nvp + v] = v < nvec ? bn[v] : 0.0;
      if (v < nvec) av[(size_t)v * nstep + i] = dd[v];
    }
    st = upload(cf);
    if (st) break;
    hipLaunchKernelGGL(k_rl_axpy, eg, dim3(256), 0, R->stream, w, (const double*)u, (const double*)up, (const double*)coef, n, nvp);
    st = coldot(w, w, dd);
    if (st) break;
    for (int v = 0; v < nvp; ++v) {
      bn[v] = v < nvec ? std::sqrt(dd[v]) : 0.0;
      if (v < nvec) bv[(size_t)v * nstep + i] = bn[v];
      cf[v] = v < nvec ? 1.0 / std::max(bn[v], 1.0e-300) : 0.0;
    }
    st = upload(cf);
    if (st) break;
    hipLaunchKernelGGL(k_rl_shift, eg, dim3(256), 0, R->stream, u, up, (const double*)w, (const double*)coef, n, nvp);
  }
  if (!st) st = hipGetLastError() != hipSuccess || hipStreamSynchronize(R->stream) != hipSuccess;
  cleanup();
  if (st) return fail("asr_gpu_spectral_bounds: a launch failed");
  double lo = 1e300, hi = -1e300;
  for (int v = 0; v < nvec; ++v) {
    std::vector<double> d(av.begin() + (size_t)v * nstep, av.begin() + (size_t)(v + 1) * nstep);
    std::vector<double> e(nstep, 0.0);
    for (int i = 0; i + 1 < nstep; ++i) e[i] = bv[(size_t)v * nstep + i];
    if (asr_host_tridiag_eigvals(nstep, d.data(), e.data())) return fail("asr_gpu_spectral_bounds: tridiagonal eigenvalues");
    lo = std::min(lo, d[0]);
    hi = std::max(hi, d[nstep - 1]);
  }
  *emin = lo - margin;
  *emax = hi + margin;
  if (a_out) std::memcpy(a_out, av.data(), av.size() * 8);
  if (b_out) std::memcpy(b_out, bv.data(), bv.size() * 8);
  return 0;
}

}  // extern "C"
