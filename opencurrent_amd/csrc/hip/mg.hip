// Multigrid kernels of Sol_MultigridPressure3DDevice<T>
// (ocuequation/sol_mgpressure3ddev.cu, sol_mgpressure3ddevd.cu, ...devf.cu).
//
// Layout reminder: k (the reference's z) has unit stride, so a wave of 64
// lanes always walks 64 consecutive k of one (i,j) row.  All arithmetic
// keeps the reference's operation order; the library is built with
// -ffp-contract=off so the results match the CPU restatement bit for bit.
#include <algorithm>
#include <climits>
#include <cstring>
#include <mutex>

#include "ghost_l1.h"
#include "rows.h"
#include "ocu_hip_common.h"

using namespace ocu_hip;

namespace {

// ------------------------------------------------------------------------
// relaxation constants, rounded exactly as each reference launcher does
// ------------------------------------------------------------------------

// Sol_MultigridPressure3DDeviceD_relax_DIAG launcher (devd.cu:1246-1268)
struct DevC64 {
  double fx, fy, fz, omega, ld, neg_hsq;
  double m[6];
};
// Sol_MultigridPressure3DDeviceF_relax launcher (devf.cu:367-395)
struct DevC32 {
  float fx, fy, fz, omega, ld, hsq;
  float m[6];
};
// do_cpu_solve (sol_mgpressure3ddev.cu:82-125)
template <typename T>
struct HostC {
  double fx, fy, fz;
  T omega, ld, h;
  T m[6];
};

DevC64 dev_c64(const ocu_mg_level_params *p) {
  DevC64 c;
  c.fx = p->fx; c.fy = p->fy; c.fz = p->fz; c.omega = p->omega;
  c.ld = (double)(2 * p->fx + 2 * p->fy + 2 * p->fz);
  c.neg_hsq = (double)-p->h * p->h;
  for (int q = 0; q < 6; q++) c.m[q] = p->mods[q];
  return c;
}
DevC32 dev_c32(const ocu_mg_level_params *p) {
  DevC32 c;
  c.fx = (float)p->fx; c.fy = (float)p->fy; c.fz = (float)p->fz; c.omega = (float)p->omega;
  c.ld = (float)(2 * p->fx + 2 * p->fy + 2 * p->fz);
  c.hsq = (float)(p->h * p->h);
  for (int q = 0; q < 6; q++) c.m[q] = (float)p->mods[q];
  return c;
}
template <typename T>
HostC<T> host_c(const ocu_mg_level_params *p) {
  HostC<T> c;
  c.fx = p->fx; c.fy = p->fy; c.fz = p->fz;
  c.omega = (T)p->omega;
  c.ld = 2 * p->fx + 2 * p->fy + 2 * p->fz;
  c.h = (T)p->h;
  for (int q = 0; q < 6; q++) c.m[q] = (T)p->mods[q];
  return c;
}

template <typename C> struct DevCFor;
template <> struct DevCFor<double> { typedef DevC64 type; };
template <> struct DevCFor<float> { typedef DevC32 type; };
template <typename T> typename DevCFor<T>::type dev_c(const ocu_mg_level_params *p);
template <> DevC64 dev_c<double>(const ocu_mg_level_params *p) { return dev_c64(p); }
template <> DevC32 dev_c<float>(const ocu_mg_level_params *p) { return dev_c32(p); }

// diagonal with boundary modifications, in the order of the diag table
// (sol_mgpressure3ddev.cu:27-43) followed by the z terms (devd.cu:585-587)
template <typename T, typename C>
__device__ __forceinline__ T diag_dev(const C &c, int i, int j, int k, int nx, int ny, int nz) {
  T d = c.ld;
  if (i == nx - 1) d += c.m[0];
  if (i == 0) d += c.m[1];
  if (j == ny - 1) d += c.m[2];
  if (j == 0) d += c.m[3];
  if (k == nz - 1) d += c.m[4];
  if (k == 0) d += c.m[5];
  return d;
}

// diag_dev in two steps: the row part (i and j faces) and the z ends on top,
// the same additions in the same order
template <typename T, typename C>
__device__ __forceinline__ T diag_row(const C &c, int i, int j, int nx, int ny) {
  T d = c.ld;
  if (i == nx - 1) d += c.m[0];
  if (i == 0) d += c.m[1];
  if (j == ny - 1) d += c.m[2];
  if (j == 0) d += c.m[3];
  return d;
}

// a uniform value held in scalar registers from here on (a kernel-argument
// load would otherwise be repeated inside every branch that uses it)
__device__ __forceinline__ float held(float x) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, x)));
}
__device__ __forceinline__ double held(double x) {
  const long long b = __builtin_bit_cast(long long, x);
  const int lo = __builtin_amdgcn_readfirstlane((int)b), hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __builtin_bit_cast(double, (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
// relax_DIAG point update (devd.cu:572-591)
__device__ __forceinline__ double relax_pt(const DevC64 &c, double uc, double zm, double zp, double ym,
                                           double yp, double xm, double xp, double b, double diag) {
  double r = -c.ld * uc;
  r += c.fz * (zm + zp);
  r += c.fy * (ym + yp);
  r += c.fx * (xm + xp);
  r += c.neg_hsq * b;
  return uc + c.omega * r / diag;
}
// Sol_MultigridPressure3DDeviceF_relax point update (devf.cu:221-243)
__device__ __forceinline__ float relax_pt(const DevC32 &c, float uc, float zm, float zp, float ym,
                                          float yp, float xm, float xp, float b, float diag) {
  float r = (-c.hsq * b - c.ld * uc);
  r += c.fz * (zm + zp);
  r += c.fy * (ym + yp);
  r += c.fx * (xm + xp);
  return uc + c.omega * r / diag;
}
// do_cpu_solve point update (sol_mgpressure3ddev.cu:106-113): neighbour
// products in double, each partial sum rounded back to T
template <typename T>
__device__ __forceinline__ T relax_pt_host_d(const HostC<T> &c, T uc, T zm, T zp, T ym, T yp, T xm, T xp, T b,
                                             T diag) {
  T r = (-c.h * c.h * b - c.ld * uc);
  r += c.fz * (zm + zp);
  r += c.fy * (yp + ym);
  r += c.fx * (xp + xm);
  return uc + c.omega * r / diag;
}
template <typename T>
__device__ __forceinline__ T diag_host(const HostC<T> &c, T x_mod, T y_mod, T z_mod) {
  return c.ld + x_mod + y_mod + z_mod;
}
template <typename T>
__device__ __forceinline__ T relax_pt_host(const HostC<T> &c, T uc, T zm, T zp, T ym, T yp, T xm, T xp,
                                           T b, T x_mod, T y_mod, T z_mod) {
  return relax_pt_host_d(c, uc, zm, zp, ym, yp, xm, xp, b, diag_host(c, x_mod, y_mod, z_mod));
}

// ------------------------------------------------------------------------
// one colour, in place
// ------------------------------------------------------------------------

// One thread per cell of the (k,j) tile; cells of the other colour idle so
// every row is read with full 64-lane coalescing.  The device variants only
// update k < 2*(nz/2) (the reference launches nz/2 threads along k).
// wg: also write the level-1 ghosts that derive from each updated cell
// (GhostP), which equals running the level-1 fill after the colour when the
// ghosts were filled before it and every periodic extent is even: a
// non-periodic face ghost is read only by its own source cell, and across an
// even periodic axis the copies of cells 0 and n-1 are read by cells of the
// other colour.
template <typename T>
__global__ __launch_bounds__(256) void k_relax_dev(DGrid<T> u, DGrid<T> b, typename DevCFor<T>::type c,
                                                   int color, GhostP<T> gp, int wg) {
  int k = blockIdx.x * TK + threadIdx.x;
  int j = blockIdx.y * TJ + threadIdx.y;
  int i = blockIdx.z;
  if (j >= u.ny || k >= 2 * (u.nz >> 1)) return;
  if (((i + j + k) & 1) != color) return;
  int64_t id = u.idx(i, j, k);
  T *p = u.p;
  T uc = p[id];
  T d = diag_dev<T>(c, i, j, k, u.nx, u.ny, u.nz);
  const T v = relax_pt(c, uc, p[id - 1], p[id + 1], p[id - u.ys], p[id + u.ys], p[id - u.xs], p[id + u.xs],
                       b.p[b.idx(i, j, k)], d);
  p[id] = v;
  if (wg) gp.write(u, i, j, k, v);
}

template <typename T>
__global__ __launch_bounds__(256) void k_relax_host(DGrid<T> u, DGrid<T> b, HostC<T> c, int color) {
  int k = blockIdx.x * TK + threadIdx.x;
  int j = blockIdx.y * TJ + threadIdx.y;
  int i = blockIdx.z;
  if (j >= u.ny || k >= u.nz) return;
  if (((i + j + k) & 1) != color) return;
  const int nx = u.nx, ny = u.ny, nz = u.nz;
  T xm_ = (i == 0) ? c.m[1] : (i == nx - 1) ? c.m[0] : (T)0;
  T ym_ = (j == 0) ? c.m[3] : (j == ny - 1) ? c.m[2] : (T)0;
  T zm_ = (k == 0) ? c.m[5] : (k == nz - 1) ? c.m[4] : (T)0;
  int64_t id = u.idx(i, j, k);
  T *p = u.p;
  p[id] = relax_pt_host(c, p[id], p[id - 1], p[id + 1], p[id - u.ys], p[id + u.ys], p[id - u.xs],
                        p[id + u.xs], b.p[b.idx(i, j, k)], xm_, ym_, zm_);
}

// ------------------------------------------------------------------------
// fused red+black iteration (out of place)
// ------------------------------------------------------------------------
//
// A block owns an output tile of OK x OJ (k,j) columns and marches along i
// over a chunk of planes.  Its 64 x 16 threads cover the tile plus a ring
// of two columns: plane p's first colour (U1) is recomputed on the tile and
// ring 1 from the old values U0 (planes p-1, p, p+1), then the second
// colour of plane p-1 is computed on the tile from U1 (planes p-2, p-1, p).
// The old values come only from u_in, so blocks never depend on each other
// and the halo recomputation gives exactly the two-launch result.
// Lateral neighbours go through LDS; the i-neighbours live in registers.

constexpr int FBK = 64, FBJ = 16, FOK = FBK - 4, FOJ = FBJ - 4, FXCD = 8;
static_assert(NORM_PARTS % FXCD == 0, "norm launches deal NORM_PARTS blocks over the XCDs");

template <typename T, bool ZERO>
__global__ __launch_bounds__(FBK *FBJ) void k_relax_fused(DGrid<T> uin, DGrid<T> uout, DGrid<T> b,
                                                          typename DevCFor<T>::type c, int first_color,
                                                          int tiles_k, int tiles_j, int ichunk, int ntiles,
                                                          GhostL1<T> gw, int wghost) {
  __shared__ T sU0[FBJ][FBK];
  __shared__ T sU1[2][FBJ][FBK];

  // blocks are dealt round-robin to the 8 XCDs: give each XCD a contiguous
  // run of tiles so neighbouring tiles share its L2
  int per = gridDim.x / FXCD;
  int logical = (blockIdx.x % FXCD) * per + (blockIdx.x / FXCD);
  if (logical >= ntiles) return;  // uniform per block
  int tk = logical % tiles_k;
  int rem = logical / tiles_k;
  int tj = rem % tiles_j;
  int ti = rem / tiles_j;

  const int tx = threadIdx.x, ty = threadIdx.y;
  const int nx = uin.nx, ny = uin.ny, nz = uin.nz;
  const int k = tk * FOK - 2 + tx;
  const int j = tj * FOJ - 2 + ty;
  const int i0 = ti * ichunk;
  const int i1 = min(nx, i0 + ichunk);
  const bool ring1 = tx >= 1 && tx <= FBK - 2 && ty >= 1 && ty <= FBJ - 2;
  const bool tile = tx >= 2 && tx <= FBK - 3 && ty >= 2 && ty <= FBJ - 3;
  const bool loadable = j >= -1 && j <= ny && k >= -1 && k <= nz;
  const bool inner_jk = j >= 0 && j < ny && k >= 0 && k < nz;
  const bool upd_k = k < 2 * (nz >> 1);
  const int64_t col = (int64_t)j * uin.ys + k;
  const int64_t bcol = (int64_t)j * b.ys + k;

  auto ld0 = [&](int p) -> T {
    if (!(loadable && p >= -1 && p <= nx)) return (T)0;
    if (ZERO) return (T)0;  // homogeneous faces: the zero grid's ghosts are +0.0
    return uin.p[(int64_t)p * uin.xs + col];
  };

  T a = ld0(i0 - 2), cc = ld0(i0 - 1);
  T u1m2 = 0, u1m1 = 0, bprev = 0;
  for (int p = i0 - 1; p <= i1; ++p) {
    T n = ld0(p + 1);
    const bool inner_p = inner_jk && p >= 0 && p < nx;
    T bcur = (ring1 && inner_p) ? b.p[(int64_t)p * b.xs + bcol] : (T)0;
    sU0[ty][tx] = cc;
    __syncthreads();
    T u1 = cc;
    if (ring1 && inner_p && upd_k && (((p + j + k) & 1) == first_color)) {
      T d = diag_dev<T>(c, p, j, k, nx, ny, nz);
      u1 = relax_pt(c, cc, sU0[ty][tx - 1], sU0[ty][tx + 1], sU0[ty - 1][tx], sU0[ty + 1][tx], a, n, bcur, d);
    }
    sU1[p & 1][ty][tx] = u1;
    __syncthreads();
    const int q = p - 1;
    if (tile && q >= i0 && inner_jk && q < nx) {
      T v = u1m1;
      if (upd_k && (((q + j + k) & 1) != first_color)) {
        T d = diag_dev<T>(c, q, j, k, nx, ny, nz);
        const int s = q & 1;
        v = relax_pt(c, u1m1, sU1[s][ty][tx - 1], sU1[s][ty][tx + 1], sU1[s][ty - 1][tx], sU1[s][ty + 1][tx],
                     u1m2, u1, bprev, d);
      }
      uout.p[(int64_t)q * uout.xs + col] = v;
      if (wghost) gw.write(uout, q, j, k, v);
    }
    a = cc; cc = n;
    u1m2 = u1m1; u1m1 = u1;
    bprev = bcur;
  }
}

// prolongation weights 27/64, 9/64, 3/64, 1/64 (devd.cu:118-223)
template <typename T> struct PW;
template <> struct PW<double> {
  static constexpr double c0 = 0.421875, c1 = 0.140625, c2 = 0.046875, c3 = 0.015625;
};
template <> struct PW<float> {
  static constexpr float c0 = 0.421875f, c1 = 0.140625f, c2 = 0.046875f, c3 = 0.015625f;
};

// a coarse row segment of VC cells per lane with its z ghost (lane 0:
// k = -1, the lane past the row: k = nzc)
template <typename T, int V>
struct CRow {
  Vec<T, V> v;
  T g;
};

// ------------------------------------------------------------------------
// fused red+black iteration, whole-row variant
// ------------------------------------------------------------------------
//
// Used when one wave holds a whole k-row: lane l owns the V consecutive
// cells k = V*l .. V*l+V-1 (V = 1, 2 or 4, the smallest with 64*V >= nz),
// so rows move as dwordx(V) accesses and no k-halo is recomputed.  The
// k-neighbours across lanes come from DPP wave shifts; the ghost cells
// k = -1 and k = nz travel in a per-lane register g (lane 0 holds k = -1;
// the lane just past the last cell, or lane 63 when the row fills the wave,
// holds k = nz).  A block of NW waves owns NW rows j (NW-4 output rows and a
// ring of two rows each side) and marches along i like the tiled kernel:
// first colour of plane p on the ring-1 rows from U0, second colour of plane
// p-1 on the output rows from U1, j-neighbours through LDS.

// One ghost row of o derived from the row (q,j) of the cells this lane wrote
// (k0 .. k0+V-1, values v): the row at offset (sx,sy) gets the x/y face
// transforms (skipped for sx = sy = 0, the row itself) and its two z-end
// ghosts the z transforms on top.
//
// With a periodic x (y) axis the ghost that derives from the low boundary
// plane (row) is the high ghost, plane nx (row ny), a copy, and the other
// way round: (sx, sy) name the boundary side of the cell, and the target
// and the transforms follow from the periodic flags.
template <typename T, int V>
__device__ __forceinline__ void write_ghost_row(const DGrid<T> &o, const GhostL1<T> &gw, int q, int j, int k0,
                                                const Vec<T, V> &v, int sx, int sy, int perx = 0, int pery = 0) {
  Vec<T, V> w;
  const int tx = perx ? 0 : sx, ty = pery ? 0 : sy;  // transformed axes
#pragma unroll
  for (int e = 0; e < V; e++) w.v[e] = gw.chain(tx, ty, 0, v.v[e]);
  const int dq = sx ? (perx ? -sx * o.nx : sx) : 0, dj = sy ? (pery ? -sy * o.ny : sy) : 0;
  T *row = o.p + (int64_t)(q + dq) * o.xs + (int64_t)(j + dj) * o.ys;
  if (sx || sy) *reinterpret_cast<Vec<T, V> *>(row + k0) = w;
  if (k0 == 0) row[-1] = gw.xf(5, w.v[0]);
  if (k0 + V == o.nz) row[o.nz] = gw.xf(4, w.v[V - 1]);
}

// One colour on a row held as V cells per lane (k = k0..k0+V-1): the cells
// with k%2 == kpar and k < kupd get the relax_DIAG update of cell (i, j, k)
// (diag_dev's diagonal), the others keep their value (lanes past the row
// too: one of them carries the k = nz ghost for its neighbour).  Even V: k0
// is even, so the relaxed elements are those with e%2 == kpar (P below, a
// uniform choice between two straight-line bodies).  V = 1: the cell is
// relaxed on every lane and kept where its colour is kpar.  The row is
// returned by value: stores through a reference from the two bodies would be
// merged into one with a computed element index, which sends the caller's
// row to scratch memory.
//
// drow: the row's diagonal without the z ends (diag_row of the row's i and
// j); the z ends go on top per cell, k = 0 only for e = 0 (k0 = V * lane).
template <int P, typename T, int V, typename C>
__device__ __forceinline__ Vec<T, V> relax_row_elems(const C &c, T drow, int kpar, int k0, int kupd, int nz,
                                                     const Vec<T, V> &cen, T lo, T hi, const Vec<T, V> &ym,
                                                     const Vec<T, V> &yp, const Vec<T, V> &xm, const Vec<T, V> &xp,
                                                     const Vec<T, V> &b, Vec<T, V> out) {
  constexpr int S = V == 1 ? 1 : 2;  // element stride
  constexpr int NE = V == 1 ? 1 : V / 2;
#pragma unroll
  for (int m = 0; m < NE; m++) {
    const int e = P + S * m, k = k0 + e;
    T d = drow;
    if (k == nz - 1) d += c.m[4];
    if (e == 0 && k == 0) d += c.m[5];
    const T nv = relax_pt(c, cen.v[e], e > 0 ? cen.v[e - 1] : lo, e < V - 1 ? cen.v[e + 1] : hi, ym.v[e], yp.v[e],
                          xm.v[e], xp.v[e], b.v[e], d);
    const bool keep = V == 1 ? ((k & 1) == kpar && k < kupd) : k < kupd;
    out.v[e] = keep ? nv : out.v[e];
  }
  return out;
}
template <typename T, int V, typename C>
__device__ __forceinline__ Vec<T, V> relax_row(const C &c, T drow, int kpar, int k0, int kupd, int nz,
                                               const Vec<T, V> &cen, T lo, T hi, const Vec<T, V> &ym,
                                               const Vec<T, V> &yp, const Vec<T, V> &xm, const Vec<T, V> &xp,
                                               const Vec<T, V> &b, const Vec<T, V> &out) {
  static_assert(V == 1 || V % 2 == 0, "rows hold 1 or an even number of cells per lane");
  if (V == 1 || kpar == 0)
    return relax_row_elems<0, T, V>(c, drow, kpar, k0, kupd, nz, cen, lo, hi, ym, yp, xm, xp, b, out);
  return relax_row_elems<1, T, V>(c, drow, kpar, k0, kupd, nz, cen, lo, hi, ym, yp, xm, xp, b, out);
}

// Rows of a block: NW waves of R rows each; the outer two rows on each side
// are the halo, so a block writes R*NW-4 rows.
//
// ZERO: u_in is the zero grid with its level-1 ghosts (not read; the values
// are the ghost transforms of 0).  PRO: the sweep's input is u_in plus the
// prolongation of uc (fine = 2 x coarse, V >= 2), with the level-1 ghosts
// that prolongation writes: each row is read from u_in, the coarse rows'
// contribution is added in registers, and ghost rows, planes and z ends are
// the face transforms of the prolonged boundary cells; u_in is not written.
// wghost: also write the level-1 ghost layer of u_out from each boundary
// cell.  ZERO, PRO and wghost need Neumann/Dirichlet faces.
// FP32 sweeps keep to 80 VGPRs (six waves per SIMD: two 12-wave blocks per
// CU, each hiding the other's barriers); the prolongation variant would
// otherwise take 81
#ifndef OCU_ROWS_F32_WAVES
#define OCU_ROWS_F32_WAVES 6
#endif
#ifndef OCU_ROWS_PRO_WAVES
#define OCU_ROWS_PRO_WAVES 6
#endif
template <typename T, bool PRO>
constexpr int fused_rows_min_waves() { return sizeof(T) == 4 ? (PRO ? OCU_ROWS_PRO_WAVES : OCU_ROWS_F32_WAVES) : 1; }

// block barriers per plane of the whole-row sweep (1 or 2; the macro exists
// for tuning builds)
#ifndef OCU_SWEEP_BARRIERS
#define OCU_SWEEP_BARRIERS 1
#endif

template <typename T, int V, int R, int NW, bool ZERO, bool PRO, bool PER>
__global__ __launch_bounds__(64 * NW, (fused_rows_min_waves<T, PRO>())) void k_relax_fused_rows(
    DGrid<T> uin, DGrid<T> uout, DGrid<T> b, typename DevCFor<T>::type c, int first_color, int tiles_j, int ichunk,
    int ntiles, GhostL1<T> gw, int wghost, DGrid<T> uc, int quirk, int perx, int pery) {
  typedef Vec<T, V> VT;
  constexpr int NR = R * NW;
  // U0 and U1 rows of planes p and p-1 (slot = plane & 1): with both
  // double-buffered one barrier per plane orders every LDS access - a wave
  // past the barrier of plane p+1 writes only slots that no wave still in
  // plane p reads (U0 of p+1, the coarse ring's plane ahead)
  __shared__ VT sU0[OCU_SWEEP_BARRIERS == 1 ? 2 : 1][NR][64];
  __shared__ VT sU1[2][NR][64];
  auto u0slot = [](int p) { return OCU_SWEEP_BARRIERS == 1 ? (p & 1) : 0; };

  int per = gridDim.x / FXCD;
  int logical = (blockIdx.x % FXCD) * per + (blockIdx.x / FXCD);
  if (logical >= ntiles) return;  // uniform per block
  const int tj = logical % tiles_j, ti = logical / tiles_j;

  const int lane = threadIdx.x, w = __builtin_amdgcn_readfirstlane(threadIdx.y);  // one row set per wave
  const int t0 = w * R;  // first block row of this wave
  const int nx = uin.nx, ny = uin.ny, nz = uin.nz;
  const int jb = tj * (NR - 4) - 2 + t0;  // j of the wave's first row
  const int k0 = lane * V;
  const int i0 = ti * ichunk, i1 = min(nx, i0 + ichunk);
  const bool active = k0 < nz;
  const int hi_lane = (nz == 64 * V) ? 63 : nz / V;
  const int gk = lane == 0 ? -1 : nz;
  const bool g_in_data = lane == hi_lane && nz < 64 * V;
  const int kupd = 2 * (nz >> 1);
  // Loads are unconditional so that the prefetches stay in flight across
  // iterations: addresses are clamped into the allocated ghost box (rows
  // -1..ny, planes -1..nx, cells 0..nz-1) and the values read for rows,
  // planes or lanes outside it only feed updates that are dropped.
  //
  // PER, periodic x (perx) and/or y (pery) with even counts, not with PRO:
  // planes and rows beyond the box are read from the other end instead of
  // the ghost layer, and the first colour is computed on them too, so the
  // second colour at the boundary sees the first colour's new values across
  // the periodic face, as the reference's fill between colours gives it.
  // Planes stay within -2..nx+1 and the rows that feed kept updates within
  // -2..ny+1, so one step of the wrap is enough (n >= 2); rows further out
  // (a last tile past 2*ny on tiny grids) are clamped into the box.
  static_assert(!(PER && PRO), "periodic faces are not taken with the prolongation");
  const int px = PER ? perx : 0, py = PER ? pery : 0;
  auto wrap1 = [](int a, int n) { return a < 0 ? a + n : a >= n ? a - n : a; };
  const int kc = active ? k0 : 0;
  const T *urow[R];
  const T *brow[R];
  int jw[R];  // the row's j, wrapped on a periodic y axis
#pragma unroll
  for (int r = 0; r < R; r++) {
    jw[r] = py ? min(max(wrap1(jb + r, ny), 0), ny - 1) : jb + r;
    urow[r] = uin.p + (int64_t)(py ? jw[r] : min(max(jb + r, -1), ny)) * uin.ys;
    brow[r] = b.p + (int64_t)(py ? jw[r] : min(max(jb + r, 0), ny - 1)) * b.ys + kc;
  }
  auto inner_j = [&](int r) { return jb + r >= 0 && jb + r < ny; };
  // the rows' diagonals without the z ends on the first plane, the last
  // plane and the planes between (uniform: held in scalar registers)
  T dlo[R], dhi[R], dmid[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    dlo[r] = held(diag_row<T>(c, 0, jw[r], nx, ny));
    dhi[r] = held(diag_row<T>(c, nx - 1, jw[r], nx, ny));
    dmid[r] = held(diag_row<T>(c, -1, jw[r], nx, ny));
  }
  auto drow = [&](int r, int i) { return i == 0 ? dlo[r] : i == nx - 1 ? dhi[r] : dmid[r]; };
  auto plane_u = [&](int p) { return px ? wrap1(p, nx) : min(max(p, -1), nx); };
  auto plane_b = [&](int p) { return px ? wrap1(p, nx) : min(max(p, 0), nx - 1); };
  auto ring1 = [&](int r) { return t0 + r >= 1 && t0 + r <= NR - 2; };
  auto tile = [&](int r) { return t0 + r >= 2 && t0 + r <= NR - 3; };

  // PRO: the row as the plain load would clamp it (jj, pp), its interior
  // source row (jr, pq) and the ghost offset between them.
  //
  // The coarse rows the block's fine rows need (rows cb .. cb+CR-1 of each
  // coarse plane, with their z ghosts) are staged in LDS, one coarse plane
  // every two fine planes, in a ring of four planes (slot = plane & 3): wave
  // w < CR loads coarse row cb+w of the plane the march will need two fine
  // planes ahead into registers and writes it at the top of the next
  // iteration, before the first barrier.  A fine row's prolongation then
  // reads its six coarse row segments (rows jc, jc +- 1 of planes ic, and
  // ic -+ 1 toward the fine plane) from LDS with their k-neighbours, so no
  // coarse window sits in registers and no DPP shift is needed.
  constexpr int VC = PRO ? V / 2 : 1;
  constexpr int CR = PRO ? NR / 2 + 2 : 1;  // coarse rows of a block
  constexpr int CW = PRO ? 64 * VC + 2 : 1; // a coarse row: k = -1 .. 64*VC
  static_assert(!PRO || (R == 1 && NW >= CR && (NR - 4) % 2 == 0), "prolongation staging needs a wave per coarse row");
  __shared__ T sC[PRO ? 4 : 1][CR][CW];
  const int cb = (jb - t0) / 2 - 1;  // the block's first fine row jb - t0 is even
  auto pro_rows = [&](int r, int p, int &jj, int &pp, int &jr, int &pq) {
    jj = min(max(jb + r, -1), ny);
    pp = min(max(p, -1), nx);
    jr = min(max(jj, 0), ny - 1);
    pq = min(max(pp, 0), nx - 1);
  };
  // this wave's coarse row of plane ic into registers (waves w < CR)
  auto cload = [&](int ic, Vec<T, VC> &d, T &g) {
    if constexpr (PRO) {
      const int row = min(max(cb + w, -1), uc.ny);
      const T *src = uc.p + (int64_t)min(max(ic, -1), uc.nx) * uc.xs + (int64_t)row * uc.ys;
      const int kcc = lane * VC < uc.nz ? lane * VC : 0;
      d = *reinterpret_cast<const Vec<T, VC> *>(src + kcc);
      g = src[lane == 0 ? -1 : uc.nz];
    }
  };
  auto cstore = [&](int ic, const Vec<T, VC> &d, T g) {
    if constexpr (PRO) {
      if (w < CR) {
        T *dst = &sC[ic & 3][w][0];
        if (lane * VC < uc.nz) {
#pragma unroll
          for (int m = 0; m < VC; m++) dst[1 + lane * VC + m] = d.v[m];
        }
        if (lane == 0) dst[0] = g;
        if (lane == hi_lane) dst[1 + uc.nz] = g;
      }
    }
  };
  auto pro_load = [&](int r, int p, VT &f) {
    if constexpr (PRO) {
      int jj, pp, jr, pq;
      pro_rows(r, p, jj, pp, jr, pq);
      f = *reinterpret_cast<const VT *>(uin.p + (int64_t)pq * uin.xs + (int64_t)jr * uin.ys + kc);
    }
  };
  // f += prolongation of the coarse correction at fine row (pq, jr): the
  // operations and their order are those of k_prolong_rows (and k_prolong),
  // so the sum is the same bit for bit
  auto pro_add = [&](int pq, int jr, VT &f) {
    if constexpr (PRO) {
      const int ic = pq >> 1, dx = pq & 1, dy = jr & 1;
      const int jc = jr >> 1, jc2 = jc + (dy ? 1 : -1);
      const int xc = dx ? ic + 1 : ic - 1;  // the coarse plane toward the fine plane
      // coarse cells k = lane*VC-1 .. lane*VC+VC of a row: a[m+1] is cell lane*VC+m
      auto seg = [&](int plane, int row, T (&a)[VC + 2]) {
        const T *src = &sC[plane & 3][row - cb][lane * VC];
#pragma unroll
        for (int m = 0; m < VC + 2; m++) a[m] = src[m];
      };
      T c0[VC + 2], c1[VC + 2], x0[VC + 2], x1[VC + 2], q1[VC + 2];
      seg(ic, jc, c0);
      seg(ic, jc2, c1);
      seg(xc, jc, x0);
      seg(xc, jc2, x1);
      const bool qrow = quirk && dx == 0 && dy == 1;  // its dz = 1 children take the corner from plane ic+1
      if (qrow) seg(ic + 1, jc2, q1);
#pragma unroll
      for (int e = 0; e < V; e++) {
        const int mm = e >> 1, dz = e & 1;
        const int zt = dz ? mm + 2 : mm;  // the coarse cell toward the child in z
        const T C = c0[mm + 1], Xv = x0[mm + 1], Y = c1[mm + 1];
        const T Z = c0[zt], XY = x1[mm + 1], XZ = x0[zt], YZ = c1[zt];
        const T XYZ = (qrow && dz) ? q1[zt] : x1[zt];
        f.v[e] = f.v[e] + PW<T>::c0 * C + PW<T>::c1 * (Xv + Y + Z) + PW<T>::c2 * (XY + XZ + YZ) + PW<T>::c3 * XYZ;
      }
    }
  };
  auto pro_finish = [&](int r, int p, VT f, VT &d, T &g) {
    if constexpr (PRO) {
      int jj, pp, jr, pq;
      pro_rows(r, p, jj, pp, jr, pq);
      pro_add(pq, jr, f);
      if (pp != pq || jj != jr) {
#pragma unroll
        for (int e = 0; e < V; e++) f.v[e] = gw.chain(pp - pq, jj - jr, 0, f.v[e]);
      }
      const T top = (nz == 64 * V) ? f.v[V - 1] : from_lower(f.v[V - 1], (T)0);
      g = lane == 0 ? gw.xf(5, f.v[0]) : gw.xf(4, top);
      d = f;
      if (g_in_data) d.v[0] = g;
    }
  };
  // the coarse plane whose neighbourhood a fine plane p needs
  auto pro_ic = [&](int p) { return min(max(min(max(p, -1), nx), 0), nx - 1) >> 1; };
  auto load = [&](int r, int p, VT &d, T &g) {
    if constexpr (ZERO) {  // homogeneous faces: every ghost of the zero grid is +0.0
#pragma unroll
      for (int e = 0; e < V; e++) d.v[e] = (T)0;
      g = (T)0;
    } else {
      const T *row = urow[r] + (int64_t)plane_u(p) * uin.xs;
      d = *reinterpret_cast<const VT *>(row + kc);
      g = row[gk];
      if (g_in_data) d.v[0] = g;
    }
  };
  auto loadb = [&](int r, int p) -> VT {
    return *reinterpret_cast<const VT *>(brow[r] + (int64_t)plane_b(p) * b.xs);
  };

  // per row: U0 planes p (cc) and p+1 (n, loaded during iteration p); U1
  // planes p-2, p-1, p; b planes p-1 and p (loaded during iteration p).  The
  // first colour's x-neighbours on plane p-1 are cells of the other colour,
  // which U1 holds unchanged, so U0 of plane p-1 is not kept: u1m1 stands in
  // for it (and starts as U0 of plane i0-2).  Registers stay under 80 so two
  // blocks share a CU and each hides the other's barriers.
  //
  // The loads run two planes ahead: iteration p issues u of plane p+2 and b
  // of plane p+1 and computes with the planes loaded one iteration earlier,
  // so a wave waits on memory only when an iteration is shorter than the
  // load latency (with the loads one plane ahead every iteration waited for
  // a full round trip).  The prefetch past the march's last plane (i1+1
  // for u, i1 for b) re-reads that plane.  PRO: the raw rows of plane p+2
  // and its coarse plane are in flight during iteration p; the row is
  // completed (prolongation added, ghost transforms) after the first
  // barrier of iteration p+1, from LDS.
#ifndef OCU_SWEEP_AHEAD
#define OCU_SWEEP_AHEAD 2
#endif
  // loads of the sweeps without the prolongation run this many planes ahead
  // (1 or 2; the macro exists for tuning builds)
  constexpr bool AHEAD2 = PRO || OCU_SWEEP_AHEAD == 2;
  VT cc[R], n[R], nn[R], u1m2[R], u1m1[R], u1[R], bprev[R], bcur[R], bnext[R];
  T ga[R], gc[R], gn[R];
  [[maybe_unused]] T gnn[R];
  [[maybe_unused]] Vec<T, VC> cpend;  // PRO: the coarse row loaded last iteration
  [[maybe_unused]] T cpend_g = 0;
  [[maybe_unused]] int clast = 0, cpend_ic = INT_MIN;  // last coarse plane loaded / waiting for its store
  if constexpr (PRO) {
    // coarse planes for fine planes i0-2 .. i0 (at most four), then the rows
    // of those three planes
    const int lo = pro_ic(i0 - 2) - 1;
    clast = pro_ic(i0) + 1;
    for (int ic = lo; ic <= clast; ++ic) {
      Vec<T, VC> d;
      T g;
      cload(ic, d, g);
      cstore(ic, d, g);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; r++) {
      VT f;
      pro_load(r, i0 - 2, f);
      pro_finish(r, i0 - 2, f, u1m1[r], ga[r]);
      pro_load(r, i0 - 1, f);
      pro_finish(r, i0 - 1, f, cc[r], gc[r]);
      pro_load(r, i0, nn[r]);  // raw; completed in the first iteration
    }
  }
#pragma unroll
  for (int r = 0; r < R; r++) {
    if constexpr (!PRO) {
      load(r, i0 - 2, u1m1[r], ga[r]);
      load(r, i0 - 1, cc[r], gc[r]);
      if constexpr (AHEAD2) load(r, i0, n[r], gn[r]);
    }
    if constexpr (AHEAD2) bcur[r] = loadb(r, i0 - 1);
#pragma unroll
    for (int e = 0; e < V; e++) u1m2[r].v[e] = bprev[r].v[e] = (T)0;
  }

  for (int p = i0 - 1; p <= i1; ++p) {
    const bool p_in = p >= 0 && p < nx;
    const int pa = min(p + 2, i1 + 1);  // the plane loaded ahead
    if constexpr (PRO) {
      if (cpend_ic != INT_MIN) cstore(cpend_ic, cpend, cpend_g);
      cpend_ic = INT_MIN;
      if (pro_ic(pa) + 1 > clast) {  // uniform: a new coarse plane every other fine plane
        clast = pro_ic(pa) + 1;
        cload(clast, cpend, cpend_g);
        cpend_ic = clast;
      }
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
      if constexpr (PRO) {
        VT raw = nn[r];  // plane p+1, completed after the barrier
        pro_load(r, pa, nn[r]);
        n[r] = raw;
      } else if constexpr (AHEAD2) {
        load(r, pa, nn[r], gnn[r]);
      } else {
        load(r, p + 1, n[r], gn[r]);
      }
      if constexpr (AHEAD2) bnext[r] = loadb(r, min(p + 1, i1));
      else bcur[r] = loadb(r, p);
      sU0[u0slot(p)][t0 + r][lane] = cc[r];
    }
    __syncthreads();
    if constexpr (PRO) {
#pragma unroll
      for (int r = 0; r < R; r++) pro_finish(r, p + 1, n[r], n[r], gn[r]);
    }

    // first colour on plane p (rows of ring 1); j-neighbours inside the
    // wave come from registers, across waves from LDS
#pragma unroll
    for (int r = 0; r < R; r++) {
      const T lo = from_lower(cc[r].v[V - 1], gc[r]), hi = from_upper(cc[r].v[0], gc[r]);
      u1[r] = cc[r];
      if (ring1(r) && (inner_j(r) || py) && (p_in || px)) {
        const int j = jw[r], pw = px ? wrap1(p, nx) : p;
        const VT ym = r > 0 ? cc[r > 0 ? r - 1 : 0] : sU0[u0slot(p)][t0 - 1][lane];
        const VT yp = r < R - 1 ? cc[r < R - 1 ? r + 1 : 0] : sU0[u0slot(p)][t0 + R][lane];
        u1[r] = relax_row<T, V>(c, drow(r, pw), (first_color ^ pw ^ j) & 1, k0, kupd, nz, cc[r], lo, hi, ym, yp,
                                u1m1[r], n[r], bcur[r], u1[r]);
      }
      sU1[p & 1][t0 + r][lane] = u1[r];
    }
    // the second colour reads U1 of plane p-1, written before this plane's
    // barrier; a second barrier is needed only when U0 has a single slot
    if (OCU_SWEEP_BARRIERS != 1) __syncthreads();

    // second colour on plane q = p-1 (output rows); U1 == U0 on ghosts
    const int q = p - 1;
    const bool q_out = q >= i0 && q < nx;
#pragma unroll
    for (int r = 0; r < R; r++) {
      const T lo = from_lower(u1m1[r].v[V - 1], ga[r]), hi = from_upper(u1m1[r].v[0], ga[r]);
      if (tile(r) && inner_j(r) && q_out) {
        const int j = jb + r;
        const VT ym = r > 0 ? u1m1[r > 0 ? r - 1 : 0] : sU1[q & 1][t0 - 1][lane];
        const VT yp = r < R - 1 ? u1m1[r < R - 1 ? r + 1 : 0] : sU1[q & 1][t0 + R][lane];
        const VT v = relax_row<T, V>(c, drow(r, q), (first_color ^ q ^ j ^ 1) & 1, k0, kupd, nz, u1m1[r], lo, hi, ym,
                                     yp, u1m2[r], u1[r], bprev[r], u1m1[r]);
        if (active) {
          *reinterpret_cast<VT *>(uout.p + (int64_t)q * uout.xs + (int64_t)j * uout.ys + k0) = v;
          if (wghost) {  // ghosts of this row: z ends, x and y face rows
            write_ghost_row<T, V>(uout, gw, q, j, k0, v, 0, 0);
            if (j == 0) write_ghost_row<T, V>(uout, gw, q, j, k0, v, 0, -1, px, py);
            if (j == ny - 1) write_ghost_row<T, V>(uout, gw, q, j, k0, v, 0, 1, px, py);
            if (q == 0) write_ghost_row<T, V>(uout, gw, q, j, k0, v, -1, 0, px, py);
            if (q == nx - 1) write_ghost_row<T, V>(uout, gw, q, j, k0, v, 1, 0, px, py);
          }
        }
      }
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
      ga[r] = gc[r];
      cc[r] = n[r]; gc[r] = gn[r];
      u1m2[r] = u1m1[r]; u1m1[r] = u1[r];
      bprev[r] = bcur[r];
      if constexpr (AHEAD2) bcur[r] = bnext[r];
      if constexpr (!PRO && AHEAD2) {
        n[r] = nn[r]; gn[r] = gnn[r];
      }
    }
  }
  if (!wghost || !active || (i0 != 0 && i1 != nx)) return;
  // the x-y edge ghost rows (boundary plane and boundary row) from this
  // wave's own stores: same lane and addresses, so program order makes the
  // values visible
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int j = jb + r;
    if (!(tile(r) && inner_j(r)) || (j != 0 && j != ny - 1)) continue;
    for (int sx = -1; sx <= 1; sx += 2) {
      const int q = sx < 0 ? 0 : nx - 1;
      if (q < i0 || q >= i1) continue;
      const VT v = *reinterpret_cast<const VT *>(uout.p + (int64_t)q * uout.xs + (int64_t)j * uout.ys + k0);
      if (j == 0) write_ghost_row<T, V>(uout, gw, q, j, k0, v, sx, -1, px, py);
      if (j == ny - 1) write_ghost_row<T, V>(uout, gw, q, j, k0, v, sx, 1, px, py);
    }
  }
}

// launch shape (rows per wave R, waves per block NW) of the whole-row kernel
// for FP32 rows and FP64 rows of four cells per lane; the macros exist for
// tuning builds
#ifndef OCU_ROWS_F32
#define OCU_ROWS_F32 1, 12
#endif
#ifndef OCU_ROWS_F64
#define OCU_ROWS_F64 1, 12
#endif
template <int R_, int NW_> struct Shape { static constexpr int R = R_, NW = NW_; };
template <typename T, int V, bool PRO> struct RowsShape : Shape<OCU_ROWS_F32> {};
template <> struct RowsShape<double, 4, false> : Shape<OCU_ROWS_F64> {};
// the sweep that adds the prolongation stages a coarse row per wave
template <> struct RowsShape<float, 4, true> : Shape<1, 12> {};
template <> struct RowsShape<double, 4, true> : Shape<1, 12> {};


// ------------------------------------------------------------------------
// residual, restriction, fused residual->restriction, norms
// ------------------------------------------------------------------------

template <typename T>
struct ResC {
  T fx, fy, fz, diag;  // f/h^2 and 2(fx+fy+fz)/h^2, rounded to T (devd.cu:1396-1399)
};
template <typename T>
ResC<T> res_c(double h, double fx, double fy, double fz) {
  ResC<T> r;
  r.fx = (T)(fx / (h * h)); r.fy = (T)(fy / (h * h)); r.fz = (T)(fz / (h * h));
  r.diag = (T)(2 * (fx + fy + fz) / (h * h));
  return r;
}

// calculate_residual point (devd.cu:108-112)
template <typename T>
__device__ __forceinline__ T res_pt(const ResC<T> &c, const T *u, int64_t id, int64_t ys, int64_t xs, T b) {
  return b - (c.fz * (u[id - 1] + u[id + 1]) + c.fy * (u[id - ys] + u[id + ys]) +
              c.fx * (u[id - xs] + u[id + xs]) - c.diag * u[id]);
}

template <typename T>
__global__ __launch_bounds__(256) void k_residual(DGrid<T> u, DGrid<T> b, DGrid<T> r, ResC<T> c) {
  int k = blockIdx.x * TK + threadIdx.x;
  int j = blockIdx.y * TJ + threadIdx.y;
  int i = blockIdx.z;
  if (j >= u.ny || k >= u.nz) return;
  r.p[r.idx(i, j, k)] = res_pt(c, u.p, u.idx(i, j, k), u.ys, u.xs, b.p[b.idx(i, j, k)]);
}

// restrict (devd.cu:65-91): child order (0,0,0),(0,0,1),(0,1,0),(0,1,1),
// (1,0,0),(1,0,1),(1,1,0),(1,1,1); `.125 * sum` with a double literal
template <typename T>
__device__ __forceinline__ T restrict_sum(T r000, T r001, T r010, T r011, T r100, T r101, T r110, T r111) {
  T s = r000 + r001 + r010 + r011 + r100 + r101 + r110 + r111;
  return (T)(.125 * s);
}

template <typename T>
__global__ __launch_bounds__(256) void k_restrict(DGrid<T> r, DGrid<T> bc) {
  int k = blockIdx.x * TK + threadIdx.x;
  int j = blockIdx.y * TJ + threadIdx.y;
  int i = blockIdx.z;
  if (j >= bc.ny || k >= bc.nz) return;
  int64_t id = r.idx(2 * i, 2 * j, 2 * k);
  const T *p = r.p;
  int64_t ys = r.ys, xs = r.xs;
  bc.p[bc.idx(i, j, k)] = restrict_sum(p[id], p[id + 1], p[id + ys], p[id + 1 + ys], p[id + xs], p[id + 1 + xs],
                                       p[id + ys + xs], p[id + 1 + ys + xs]);
}

// Fused: each thread owns a coarse cell, computes its 8 fine residuals and
// their restriction.  With NORM the grid is capped at NORM_PARTS blocks
// that loop over tiles and leave (sum r^2, max|r|) partials.
template <typename T, bool NORM>
__global__ __launch_bounds__(256) void k_res_restrict(DGrid<T> u, DGrid<T> b, DGrid<T> r, bool write_r,
                                                      DGrid<T> bc, ResC<T> c, int tiles_k, int tiles_j,
                                                      int ntiles, double *parts) {
  double s2 = 0, mx = 0;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    int tk = t % tiles_k, rem = t / tiles_k, tj = rem % tiles_j, i = rem / tiles_j;
    int k = tk * TK + threadIdx.x;
    int j = tj * TJ + threadIdx.y;
    if (j >= bc.ny || k >= bc.nz) continue;
    T rr[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      int di = q >> 2, dj = (q >> 1) & 1, dk = q & 1;
      int I = 2 * i + di, J = 2 * j + dj, K = 2 * k + dk;
      rr[q] = res_pt(c, u.p, u.idx(I, J, K), u.ys, u.xs, b.p[b.idx(I, J, K)]);
      if (write_r) r.p[r.idx(I, J, K)] = rr[q];
      if (NORM) {
        double x = (double)rr[q];
        s2 += x * x;
        mx = fmax(mx, fabs(x));
      }
    }
    bc.p[bc.idx(i, j, k)] = restrict_sum(rr[0], rr[1], rr[2], rr[3], rr[4], rr[5], rr[6], rr[7]);
  }
  if (NORM) {
    __shared__ double sh[2][4];
    int tid = threadIdx.y * blockDim.x + threadIdx.x;
    s2 = wave_sum(s2);
    mx = wave_max(mx);
    if ((tid & 63) == 0) { sh[0][tid >> 6] = s2; sh[1][tid >> 6] = mx; }
    __syncthreads();
    if (tid == 0) {
      double a = 0, m = 0;
      for (int w = 0; w < 4; w++) { a += sh[0][w]; m = fmax(m, sh[1][w]); }
      parts[2 * blockIdx.x] = a;
      parts[2 * blockIdx.x + 1] = m;
    }
  }
}

// Residual + restriction with whole rows: each wave owns one coarse row jc,
// i.e. the fine rows 2jc and 2jc+1 held as V fine cells per lane (V/2
// coarse cells), and marches along the fine planes of its i-chunk.  The
// fine planes P-1, P, P+1 of the two rows stay in registers, the j-halo rows
// 2jc-1 and 2jc+2 of plane P are loaded beside them (L2 hits: they are
// other waves' rows), k-neighbours come by DPP with the z ghosts of the row
// in g (lane 0: k = -1; the lane after the row, or lane 63: k = nz).  No LDS,
// no barriers; operation orders are those of res_pt and restrict_sum.
template <typename T>
__device__ __forceinline__ T res_val(const ResC<T> &c, T b, T zm, T zp, T ym, T yp, T xm, T xp, T uc) {
  return b - (c.fz * (zm + zp) + c.fy * (ym + yp) + c.fx * (xm + xp) - c.diag * uc);
}

//
// NORM: the fine residuals' (sum r^2, max|r|) in double, accumulated by each
// lane in march order, reduced over the wave (fixed butterfly) and over the
// block's waves in wave order, go to parts[logical tile]; the launch has
// exactly NORM_PARTS blocks and the blocks past the last tile write zeros, so
// the partials need no clearing and the result does not depend on timing.
template <typename T, int V, int NW, bool NORM>
__global__ __launch_bounds__(64 * NW) void k_res_restrict_rows(DGrid<T> u, DGrid<T> b, DGrid<T> bc, ResC<T> c,
                                                                int tiles_j, int ichunk, int ntiles, double *parts) {
  typedef Vec<T, V> VT;
  typedef Vec<T, V / 2> VC;
  // the first nb blocks (ntiles rounded up to the XCDs) take the tiles, each
  // XCD a contiguous run of them; with norms every block of the NORM_PARTS
  // grid writes its partial pair (zeros past the tiles)
  const int nb = ((ntiles + FXCD - 1) / FXCD) * FXCD;
  const int per = nb / FXCD;
  const int logical = blockIdx.x < nb ? (blockIdx.x % FXCD) * per + (blockIdx.x / FXCD) : ntiles;
  if (logical >= ntiles) {  // uniform per block
    if (NORM && threadIdx.x == 0 && threadIdx.y == 0) {
      parts[2 * blockIdx.x] = 0.0;
      parts[2 * blockIdx.x + 1] = 0.0;
    }
    return;
  }
  const int tj = logical % tiles_j, ti = logical / tiles_j;
  const int lane = threadIdx.x, w = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int jc = tj * NW + w;
  double s2 = 0.0, mx = 0.0;
  // a wave past the last coarse row does no work (uniform per wave); without
  // NORM it leaves, with NORM it still joins the block's reduction
  if (jc < bc.ny) {
  const int nz = u.nz;
  const int k0 = lane * V;
  const bool active = k0 < nz;
  const int kc = active ? k0 : 0;
  const int hi_lane = (nz == 64 * V) ? 63 : nz / V;
  const int gk = lane == 0 ? -1 : nz;
  const bool g_in_data = lane == hi_lane && nz < 64 * V;
  const int ic0 = ti * ichunk, ic1 = min(bc.nx, ic0 + ichunk);
  // fine rows 2jc-1 .. 2jc+2 (rows -1 and ny are ghost rows)
  const T *urow[4];
#pragma unroll
  for (int q = 0; q < 4; q++) urow[q] = u.p + (int64_t)(2 * jc - 1 + q) * u.ys;
  const T *brow[2] = {b.p + (int64_t)(2 * jc) * b.ys + kc, b.p + (int64_t)(2 * jc + 1) * b.ys + kc};
  auto row = [&](int q, int P) { return *reinterpret_cast<const VT *>(urow[q] + (int64_t)P * u.xs + kc); };
  auto rowg = [&](int q, int P, VT &d, T &g) {
    const T *r = urow[q] + (int64_t)P * u.xs;
    d = *reinterpret_cast<const VT *>(r + kc);
    g = r[gk];
    if (g_in_data) d.v[0] = g;
  };

  VT prev[2], cur[2], nxt[2], rr0[2];
  T gc[2], gn[2];
#pragma unroll
  for (int q = 0; q < 2; q++) {
    prev[q] = row(q + 1, 2 * ic0 - 1);
    rowg(q + 1, 2 * ic0, cur[q], gc[q]);
  }
  for (int P = 2 * ic0; P < 2 * ic1; ++P) {
    VT h0 = row(0, P), h3 = row(3, P), bb[2];
#pragma unroll
    for (int q = 0; q < 2; q++) {
      rowg(q + 1, P + 1, nxt[q], gn[q]);
      bb[q] = *reinterpret_cast<const VT *>(brow[q] + (int64_t)P * b.xs);
    }
    VT rr[2];
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const T lo = from_lower(cur[q].v[V - 1], gc[q]), hi = from_upper(cur[q].v[0], gc[q]);
      const VT &ym = q == 0 ? h0 : cur[0];
      const VT &yp = q == 0 ? cur[1] : h3;
#pragma unroll
      for (int e = 0; e < V; e++)
        rr[q].v[e] = res_val(c, bb[q].v[e], e > 0 ? cur[q].v[e - 1] : lo, e < V - 1 ? cur[q].v[e + 1] : hi, ym.v[e],
                             yp.v[e], prev[q].v[e], nxt[q].v[e], cur[q].v[e]);
      if (NORM && active) {
#pragma unroll
        for (int e = 0; e < V; e++) {
          const double x = (double)rr[q].v[e];
          s2 += x * x;
          mx = fmax(mx, fabs(x));
        }
      }
    }
    if (P & 1) {
      VC o;
#pragma unroll
      for (int m = 0; m < V / 2; m++)
        o.v[m] = restrict_sum(rr0[0].v[2 * m], rr0[0].v[2 * m + 1], rr0[1].v[2 * m], rr0[1].v[2 * m + 1],
                              rr[0].v[2 * m], rr[0].v[2 * m + 1], rr[1].v[2 * m], rr[1].v[2 * m + 1]);
      if (active) *reinterpret_cast<VC *>(bc.p + (int64_t)(P >> 1) * bc.xs + (int64_t)jc * bc.ys + lane * (V / 2)) = o;
    } else {
      rr0[0] = rr[0];
      rr0[1] = rr[1];
    }
#pragma unroll
    for (int q = 0; q < 2; q++) {
      prev[q] = cur[q];
      cur[q] = nxt[q];
      gc[q] = gn[q];
    }
  }
  } else if (!NORM) {
    return;
  }
  if constexpr (NORM) {
    __shared__ double sh[2][NW];
    s2 = wave_sum(s2);
    mx = wave_max(mx);
    if (lane == 0) {
      sh[0][w] = s2;
      sh[1][w] = mx;
    }
    __syncthreads();
    if (w == 0 && lane == 0) {
      double a = 0.0, m = 0.0;
      for (int q = 0; q < NW; q++) {
        a += sh[0][q];
        m = fmax(m, sh[1][q]);
      }
      parts[2 * blockIdx.x] = a;
      parts[2 * blockIdx.x + 1] = m;
    }
  }
}

// residual over the fine grid with norm partials (no restriction)
template <typename T>
__global__ __launch_bounds__(256) void k_res_norm(DGrid<T> u, DGrid<T> b, DGrid<T> r, ResC<T> c, int tiles_k,
                                                  int tiles_j, int ntiles, double *parts) {
  double s2 = 0, mx = 0;
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    int tk = t % tiles_k, rem = t / tiles_k, tj = rem % tiles_j, i = rem / tiles_j;
    int k = tk * TK + threadIdx.x;
    int j = tj * TJ + threadIdx.y;
    if (j >= u.ny || k >= u.nz) continue;
    T x = res_pt(c, u.p, u.idx(i, j, k), u.ys, u.xs, b.p[b.idx(i, j, k)]);
    r.p[r.idx(i, j, k)] = x;
    double d = (double)x;
    s2 += d * d;
    mx = fmax(mx, fabs(d));
  }
  __shared__ double sh[2][4];
  int tid = threadIdx.y * blockDim.x + threadIdx.x;
  s2 = wave_sum(s2);
  mx = wave_max(mx);
  if ((tid & 63) == 0) { sh[0][tid >> 6] = s2; sh[1][tid >> 6] = mx; }
  __syncthreads();
  if (tid == 0) {
    double a = 0, m = 0;
    for (int w = 0; w < 4; w++) { a += sh[0][w]; m = fmax(m, sh[1][w]); }
    parts[2 * blockIdx.x] = a;
    parts[2 * blockIdx.x + 1] = m;
  }
}

// tag != 0: after the two results, with a system-scope fence between, the
// tag goes into out[2] (host memory that the host polls)
__global__ void k_norm_finalize(const double *parts, int n, double *out, unsigned long long tag) {
  __shared__ double sh[2][8];
  double s = 0, m = 0;
  for (int q = threadIdx.x; q < n; q += blockDim.x) {
    s += parts[2 * q];
    m = fmax(m, parts[2 * q + 1]);
  }
  s = wave_sum(s);
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) { sh[0][threadIdx.x >> 6] = s; sh[1][threadIdx.x >> 6] = m; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0, b = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) { a += sh[0][w]; b = fmax(b, sh[1][w]); }
    out[0] = a;
    out[1] = b;
    if (tag) {
      __threadfence_system();
      *reinterpret_cast<volatile unsigned long long *>(out + 2) = tag;
    }
  }
}

// ------------------------------------------------------------------------
// prolongation + correction (devd.cu:118-223, devf.cu:94-197)
// ------------------------------------------------------------------------


template <typename T>
__global__ __launch_bounds__(256) void k_prolong(DGrid<T> uc, DGrid<T> uf, int quirk, GhostL1<T> gw, int wghost) {
  int kc = blockIdx.x * TK + threadIdx.x;
  int jc = blockIdx.y * TJ + threadIdx.y;
  int ic = blockIdx.z;
  if (jc >= uc.ny || kc >= uc.nz) return;
  const T *q = uc.p;
  const int64_t xs = uc.xs, ys = uc.ys;
  const int64_t id = uc.idx(ic, jc, kc);
  const T C = q[id];
#pragma unroll
  for (int n = 0; n < 8; n++) {
    const int dx = n >> 2, dy = (n >> 1) & 1, dz = n & 1;
    const int64_t sx = dx ? xs : -xs, sy = dy ? ys : -ys, sz = dz ? 1 : -1;
    const int64_t cx = (quirk && dx == 0 && dy == 1 && dz == 1) ? xs : sx;
    T X = q[id + sx], Y = q[id + sy], Z = q[id + sz];
    T XY = q[id + sx + sy], XZ = q[id + sx + sz], YZ = q[id + sy + sz];
    T XYZ = q[id + cx + sy + sz];
    T *f = &uf.at(2 * ic + dx, 2 * jc + dy, 2 * kc + dz);
    const T v = *f + PW<T>::c0 * C + PW<T>::c1 * (X + Y + Z) + PW<T>::c2 * (XY + XZ + YZ) + PW<T>::c3 * XYZ;
    *f = v;
    if (wghost) gw.write(uf, 2 * ic + dx, 2 * jc + dy, 2 * kc + dz, v);
  }
}

// Prolongation with whole rows: each wave owns one fine row J (V fine cells
// per lane, V/2 coarse cells) and marches along the fine planes I of its
// chunk.  The coarse rows it reads - rows jc and jc+sy (sy = +-1 toward the
// fine row) of the coarse planes ic-1, ic, ic+1 - stay in registers as a
// sliding window, with their z ghosts in g for the DPP shifts; the fine row
// is read, corrected and written once.  Same operations as k_prolong,
// including its corner choice and the level-1 ghost writes.

template <typename T, int V, int NW>
__global__ __launch_bounds__(64 * NW) void k_prolong_rows(DGrid<T> uc, DGrid<T> uf, int quirk, GhostL1<T> gw,
                                                          int wghost, int tiles_j, int ichunk, int ntiles, int perx,
                                                          int pery) {
  constexpr int VC = V / 2;
  typedef Vec<T, V> VT;
  typedef CRow<T, VC> CR;
  const int per = gridDim.x / FXCD;
  const int logical = (blockIdx.x % FXCD) * per + (blockIdx.x / FXCD);
  if (logical >= ntiles) return;  // uniform per block
  const int tj = logical % tiles_j, ti = logical / tiles_j;
  const int lane = threadIdx.x, w = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int J = tj * NW + w;
  if (J >= uf.ny) return;  // uniform per wave; no barriers in this kernel
  const int nzc = uc.nz;
  const int k0 = lane * V;
  const bool active = k0 < uf.nz;
  const int kc = active ? lane * VC : 0;
  const int hi_lane = (uf.nz == 64 * V) ? 63 : uf.nz / V;
  const int gk = lane == 0 ? -1 : nzc;
  const bool g_in_data = lane == hi_lane && uf.nz < 64 * V;
  const int jc = J >> 1, dy = J & 1;
  const int I0 = 2 * (ti * ichunk), I1 = min(uf.nx, I0 + 2 * ichunk);
  // coarse rows jc (r = 0) and jc + (dy ? 1 : -1) (r = 1)
  const T *crow[2] = {uc.p + (int64_t)jc * uc.ys, uc.p + (int64_t)(jc + (dy ? 1 : -1)) * uc.ys};
  auto load = [&](int r, int ic, CR &d) {
    const T *p = crow[r] + (int64_t)ic * uc.xs;
    d.v = *reinterpret_cast<const Vec<T, VC> *>(p + kc);
    d.g = p[gk];
    if (g_in_data) d.v.v[0] = d.g;
  };
  // window: coarse planes ic-1 (m), ic (c), ic+1 (p)
  CR m[2], c[2], p[2];
  int ic = I0 >> 1;
#pragma unroll
  for (int r = 0; r < 2; r++) {
    load(r, ic - 1, m[r]);
    load(r, ic, c[r]);
    load(r, ic + 1, p[r]);
  }
  // the fine row of the next plane is read one iteration ahead (other
  // planes than the one being written), so its latency overlaps the work
  auto load_f = [&](int I) {
    return *reinterpret_cast<const VT *>(uf.p + (int64_t)I * uf.xs + (int64_t)J * uf.ys + (active ? k0 : 0));
  };
  VT fnext = load_f(I0);
  for (int I = I0; I < I1; ++I) {
    const int dx = I & 1;
    VT f = fnext;
    if (I + 1 < I1) fnext = load_f(I + 1);
    if (I > I0 && !dx) {  // entering coarse plane ic+1
      ++ic;
#pragma unroll
      for (int r = 0; r < 2; r++) {
        m[r] = c[r];
        c[r] = p[r];
        load(r, ic + 1, p[r]);
      }
    }
    const CR *X = dx ? p : m;  // the coarse plane toward the fine plane
    // coarse values at kc-1 (toward dz = 0) and kc+VC (toward dz = 1) of
    // each row: lane neighbours by DPP, the row's ghost at the ends
    T zl[2][2], zh[2][2];  // [plane: 0 = c, 1 = X][row]
#pragma unroll
    for (int r = 0; r < 2; r++) {
      zl[0][r] = from_lower(c[r].v.v[VC - 1], c[r].g);
      zh[0][r] = from_upper(c[r].v.v[0], c[r].g);
      zl[1][r] = from_lower(X[r].v.v[VC - 1], X[r].g);
      zh[1][r] = from_upper(X[r].v.v[0], X[r].g);
    }
    const bool qrow = quirk && dx == 0 && dy == 1;  // its dz = 1 children take the corner from plane ic+1
    T qzh = 0;
    if (qrow) qzh = from_upper(p[1].v.v[0], p[1].g);
    T *frow = uf.p + (int64_t)I * uf.xs + (int64_t)J * uf.ys;
#pragma unroll
    for (int e = 0; e < V; e++) {
      const int mm = e >> 1, dz = e & 1;
      auto zs = [&](const CR &row, T lo, T hi) {  // coarse value toward the child in z
        return dz ? (mm + 1 < VC ? row.v.v[mm + 1] : hi) : (mm > 0 ? row.v.v[mm - 1] : lo);
      };
      const T C = c[0].v.v[mm], Xv = X[0].v.v[mm], Y = c[1].v.v[mm];
      const T Z = zs(c[0], zl[0][0], zh[0][0]);
      const T XY = X[1].v.v[mm];
      const T XZ = zs(X[0], zl[1][0], zh[1][0]);
      const T YZ = zs(c[1], zl[0][1], zh[0][1]);
      const T XYZ = (qrow && dz) ? (mm + 1 < VC ? p[1].v.v[mm + 1] : qzh) : zs(X[1], zl[1][1], zh[1][1]);
      f.v[e] = f.v[e] + PW<T>::c0 * C + PW<T>::c1 * (Xv + Y + Z) + PW<T>::c2 * (XY + XZ + YZ) + PW<T>::c3 * XYZ;
    }
    if (active) {
      *reinterpret_cast<VT *>(frow + k0) = f;
      if (wghost) {
        const bool xl = I == 0, xh = I == uf.nx - 1, yl = J == 0, yh = J == uf.ny - 1;
        write_ghost_row<T, V>(uf, gw, I, J, k0, f, 0, 0);
        if (yl) write_ghost_row<T, V>(uf, gw, I, J, k0, f, 0, -1, perx, pery);
        if (yh) write_ghost_row<T, V>(uf, gw, I, J, k0, f, 0, 1, perx, pery);
        if (xl) write_ghost_row<T, V>(uf, gw, I, J, k0, f, -1, 0, perx, pery);
        if (xh) write_ghost_row<T, V>(uf, gw, I, J, k0, f, 1, 0, perx, pery);
        if (xl && yl) write_ghost_row<T, V>(uf, gw, I, J, k0, f, -1, -1, perx, pery);
        if (xl && yh) write_ghost_row<T, V>(uf, gw, I, J, k0, f, -1, 1, perx, pery);
        if (xh && yl) write_ghost_row<T, V>(uf, gw, I, J, k0, f, 1, -1, perx, pery);
        if (xh && yh) write_ghost_row<T, V>(uf, gw, I, J, k0, f, 1, 1, perx, pery);
      }
    }
  }
}

// ------------------------------------------------------------------------
// coarsest level in one workgroup (LDS resident)
// ------------------------------------------------------------------------

struct CoarseGeom {
  int nx, ny, nz;
  int lys, lxs;   // LDS strides of the (nx+2)(ny+2)(nz+2) box
  int lshift;     // LDS offset of cell (0,0,0)
};

// The threads that share an LDS phase (the whole block) and their barrier.
struct BlockTeam {
  __device__ int rank() const { return threadIdx.x; }
  __device__ int size() const { return blockDim.x; }
  __device__ void sync() const { __syncthreads(); }
};

template <typename T, typename Team>
__device__ void lds_relax_color(const Team &tm, T *u, const T *b, const CoarseGeom &g, const HostC<T> &c, int color) {
  int n = g.nx * g.ny * g.nz;
  for (int t = tm.rank(); t < n; t += tm.size()) {
    int k = t % g.nz, j = (t / g.nz) % g.ny, i = t / (g.nz * g.ny);
    if (((i + j + k) & 1) != color) continue;
    int id = g.lshift + i * g.lxs + j * g.lys + k;
    T xm_ = (i == 0) ? c.m[1] : (i == g.nx - 1) ? c.m[0] : (T)0;
    T ym_ = (j == 0) ? c.m[3] : (j == g.ny - 1) ? c.m[2] : (T)0;
    T zm_ = (k == 0) ? c.m[5] : (k == g.nz - 1) ? c.m[4] : (T)0;
    u[id] = relax_pt_host(c, u[id], u[id - 1], u[id + 1], u[id - g.lys], u[id + g.lys], u[id - g.lxs],
                          u[id + g.lxs], b[id], xm_, ym_, zm_);
  }
}

struct FaceS {
  int type;
  float value;
};
struct FacesS {
  FaceS f[6];  // xpos, xneg, ypos, yneg, zpos, zneg
};

__device__ __forceinline__ bool wr(const FaceS &f) {
  return f.type == OCU_BC_PERIODIC || f.type == OCU_BC_DIRICHELET || f.type == OCU_BC_NEUMANN;
}
// host no-corner ghost (grid3dboundary.cu:1018-1111): Dirichlet in T
template <typename T>
__device__ __forceinline__ T xf_host(const FaceS &f, bool neg, T h, T a) {
  if (f.type == OCU_BC_DIRICHELET) return -a + (T)2 * (T)f.value;
  if (f.type == OCU_BC_NEUMANN) return neg ? a - h * (T)f.value : a + h * (T)f.value;
  return a;
}
// level-1 ghost (grid3dboundary.cu:237-347): Dirichlet `-phi + 2 * value`
template <typename T>
__device__ __forceinline__ T xf_l1(const FaceS &f, bool neg, T h, T a) {
  if (f.type == OCU_BC_DIRICHELET) return -a + (T)(2 * f.value);
  if (f.type == OCU_BC_NEUMANN) return neg ? a - h * (T)f.value : a + h * (T)f.value;
  return a;
}
__device__ __forceinline__ int srci(const FaceS &f, bool neg, int n) {
  if (f.type == OCU_BC_PERIODIC) return neg ? n - 1 : 0;
  return neg ? 0 : n - 1;
}

template <typename T, typename Team>
__device__ void lds_fill_nocorners(const Team &tm, T *u, const CoarseGeom &g, const FacesS &F, T hx, T hy, T hz) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  auto A = [&](int i, int j, int k) -> T & { return u[g.lshift + i * g.lxs + j * g.lys + k]; };
  int nZ = nx * ny, nY = nx * nz, nX = ny * nz;
  for (int t = tm.rank(); t < nZ + nY + nX; t += tm.size()) {
    if (t < nZ) {
      int i = t / ny, j = t % ny;
      if (wr(F.f[5])) A(i, j, -1) = xf_host(F.f[5], true, hz, A(i, j, srci(F.f[5], true, nz)));
      if (wr(F.f[4])) A(i, j, nz) = xf_host(F.f[4], false, hz, A(i, j, srci(F.f[4], false, nz)));
    } else if (t < nZ + nY) {
      int q = t - nZ, i = q / nz, k = q % nz;
      if (wr(F.f[3])) A(i, -1, k) = xf_host(F.f[3], true, hy, A(i, srci(F.f[3], true, ny), k));
      if (wr(F.f[2])) A(i, ny, k) = xf_host(F.f[2], false, hy, A(i, srci(F.f[2], false, ny), k));
    } else {
      int q = t - nZ - nY, j = q / nz, k = q % nz;
      if (wr(F.f[1])) A(-1, j, k) = xf_host(F.f[1], true, hx, A(srci(F.f[1], true, nx), j, k));
      if (wr(F.f[0])) A(nx, j, k) = xf_host(F.f[0], false, hx, A(srci(F.f[0], false, nx), j, k));
    }
  }
}

// the level-1 fill, pass by pass (block barriers between passes)
template <typename T, typename Team>
__device__ void lds_fill_level1(const Team &tm, T *u, const CoarseGeom &g, const FacesS &F, T hx, T hy, T hz) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  auto A = [&](int i, int j, int k) -> T & { return u[g.lshift + i * g.lxs + j * g.lys + k]; };
  for (int t = tm.rank(); t < ny * nz; t += tm.size()) {
    int j = t / nz, k = t % nz;
    if (wr(F.f[1])) A(-1, j, k) = xf_l1(F.f[1], true, hx, A(srci(F.f[1], true, nx), j, k));
    if (wr(F.f[0])) A(nx, j, k) = xf_l1(F.f[0], false, hx, A(srci(F.f[0], false, nx), j, k));
  }
  tm.sync();
  for (int t = tm.rank(); t < (nx + 2) * nz; t += tm.size()) {
    int i = t / nz - 1, k = t % nz;
    if (wr(F.f[3])) A(i, -1, k) = xf_l1(F.f[3], true, hy, A(i, srci(F.f[3], true, ny), k));
    if (wr(F.f[2])) A(i, ny, k) = xf_l1(F.f[2], false, hy, A(i, srci(F.f[2], false, ny), k));
  }
  tm.sync();
  for (int t = tm.rank(); t < (nx + 2) * (ny + 2); t += tm.size()) {
    int i = t / (ny + 2) - 1, j = t % (ny + 2) - 1;
    if (wr(F.f[5])) A(i, j, -1) = xf_l1(F.f[5], true, hz, A(i, j, srci(F.f[5], true, nz)));
    if (wr(F.f[4])) A(i, j, nz) = xf_l1(F.f[4], false, hz, A(i, j, srci(F.f[4], false, nz)));
  }
  tm.sync();
}

// colour sequence of relax() / relax_on_host (sol_mgpressure3ddev.cu:160-171)
__device__ __forceinline__ int color_of(int order, int it, int iterations, int rb) {
  return (order == 1 || (order == 2 && it >= iterations / 2)) ? !rb : rb;
}

// TS = grid storage type, TC = compute type.  MODE 0: reference semantics;
// MODE 1: double arithmetic on float storage with level-1 fills.
// relax_on_host semantics on LDS boxes u, b (compute type TC)
template <typename TC, int MODE, typename Team>
__device__ void coarse_iterations(const Team &tm, TC *u, TC *b, const CoarseGeom &g, const HostC<TC> &c,
                                  const FacesS &F, TC hx, TC hy, TC hz, int iterations, int order) {
  for (int it = 0; it < iterations; it++)
    for (int rb = 0; rb < 2; rb++) {
      lds_relax_color(tm, u, b, g, c, color_of(order, it, iterations, rb));
      tm.sync();
      if (MODE == 0) {
        lds_fill_nocorners(tm, u, g, F, hx, hy, hz);
        tm.sync();
      } else {
        lds_fill_level1(tm, u, g, F, hx, hy, hz);
      }
    }
  if (MODE == 0) lds_fill_level1(tm, u, g, F, hx, hy, hz);
}

// TS = grid storage type, TC = compute type.  MODE 0: reference semantics;
// MODE 1: double arithmetic on float storage with level-1 fills.
template <typename TS, typename TC, int MODE>
__global__ __launch_bounds__(256) void k_coarse(DGrid<TS> ug, DGrid<TS> bg, CoarseGeom g, HostC<TC> c, FacesS F,
                                                TC hx, TC hy, TC hz, int iterations, int order) {
  extern __shared__ unsigned char smem[];
  const BlockTeam tm;
  const int box = (g.nx + 2) * (g.ny + 2) * (g.nz + 2);
  TC *u = (TC *)smem;
  TC *b = u + box;
  for (int t = tm.rank(); t < box; t += tm.size()) {
    int k = t % (g.nz + 2) - 1, j = (t / (g.nz + 2)) % (g.ny + 2) - 1, i = t / ((g.nz + 2) * (g.ny + 2)) - 1;
    u[t] = (TC)ug.at(i, j, k);
    b[t] = (TC)bg.at(i, j, k);
  }
  tm.sync();
  coarse_iterations<TC, MODE>(tm, u, b, g, c, F, hx, hy, hz, iterations, order);
  for (int t = tm.rank(); t < box; t += tm.size()) {
    int k = t % (g.nz + 2) - 1, j = (t / (g.nz + 2)) % (g.ny + 2) - 1, i = t / ((g.nz + 2) * (g.ny + 2)) - 1;
    ug.at(i, j, k) = (TS)u[t];
  }
}

// A coarsest level whose boxes exceed one workgroup's LDS (a grid that
// coarsens little or not at all): the same iterations by one workgroup on
// global memory.  MODE 0 works on the grids in place (u and b share one
// layout); MODE 1 on double boxes wu, wb of a temporary, loaded from and
// stored back to the float grids as k_coarse does with LDS.  A workgroup
// runs on one CU, so its barriers order its global accesses as they order
// LDS ones.
template <typename TS, typename TC, int MODE>
__global__ __launch_bounds__(1024) void k_coarse_global(DGrid<TS> ug, DGrid<TS> bg, TC *wu, TC *wb, CoarseGeom g,
                                                        HostC<TC> c, FacesS F, TC hx, TC hy, TC hz, int iterations,
                                                        int order) {
  const BlockTeam tm;
  const int box = (g.nx + 2) * (g.ny + 2) * (g.nz + 2);
  if (MODE == 1) {
    for (int t = tm.rank(); t < box; t += tm.size()) {
      int k = t % (g.nz + 2) - 1, j = (t / (g.nz + 2)) % (g.ny + 2) - 1, i = t / ((g.nz + 2) * (g.ny + 2)) - 1;
      wu[t] = (TC)ug.at(i, j, k);
      wb[t] = (TC)bg.at(i, j, k);
    }
    tm.sync();
  }
  coarse_iterations<TC, MODE>(tm, wu, wb, g, c, F, hx, hy, hz, iterations, order);
  if (MODE == 1) {
    tm.sync();
    for (int t = tm.rank(); t < box; t += tm.size()) {
      int k = t % (g.nz + 2) - 1, j = (t / (g.nz + 2)) % (g.ny + 2) - 1, i = t / ((g.nz + 2) * (g.ny + 2)) - 1;
      ug.at(i, j, k) = (TS)wu[t];
    }
  }
}

// ------------------------------------------------------------------------
// the bottom of a V-cycle in one workgroup (LDS resident)
// ------------------------------------------------------------------------
//
// Levels top..C (C the coarsest) are small enough to live in LDS together;
// one launch performs the V-cycle's part on them exactly as the per-level
// launches do - relax(nu1) with the device arithmetic, a zero start on the
// next level, fused residual and restriction, the coarsest relaxation
// (ocu_mg_coarse_relax semantics), prolongation with the level-1 fill,
// relax(nu2) - and writes every level back.
//
// Ghosts: the 7-point stencil reads a face ghost only from the cell it
// derives from, so a colour pass writes each updated cell's face ghosts
// right away (no fill phase between colours or iterations); the edges and
// corners, which no relaxation reads, are filled once at the end of each
// relax call.  With Neumann/Dirichlet faces every variant of the reference's
// fills (level-1, device and host no-corner) gives these face values, so the
// final state is the reference's.

constexpr int BOT_MAX = 6;  // levels in one bottom launch
constexpr int BOT_THREADS = 1024;

// x / n for 0 <= x < 2^16 and 1 <= n <= 2^15 by one multiply: with
// m = ceil(2^31 / n) the error of x*m/2^31 is below x/2^31 <= 1/n
struct Mag {
  uint32_t n, m;
};
inline Mag mag(uint32_t n) { return {n, (uint32_t)(((1ull << 31) + n - 1) / n)}; }
__device__ __forceinline__ uint32_t mdiv(uint32_t x, Mag d) { return (uint32_t)(((uint64_t)x * d.m) >> 31); }

struct BotGeom {
  int nx, ny, nz;
  int lys, lxs, lshift;  // LDS box strides and the offset of cell (0,0,0)
  Mag my, mz, mpy, mpz;  // ny, nz, ny+2, nz+2
  Mag mhd, mhh, mn;      // kupd/2 (device colour slots per row), ceil(nz/2), nx*ny*nz
};

template <typename T>
struct BotLevel {
  DGrid<T> u, b;  // global grids of the level
  BotGeom g;
  int off;  // LDS byte offset of this level's u box (its b box follows)
  typename DevCFor<T>::type dc;
  ResC<T> rc;
  GhostL1<T> gw;
};

// periodic axes of the bottom (x, y, z): a ghost there is a copy of the
// interior cell at the other end instead of a face transform
struct BotPer {
  int a[3];
};

template <typename T, typename TC>
struct BotArgs {
  BotLevel<T> L[BOT_MAX];
  BotPer per;
  int nlev, nu1, nu2, zero_top;
  HostC<TC> hc;     // coarsest level: relax_on_host arithmetic in TC
  GhostL1<TC> gwc;  // and its ghost transforms in TC
  int coarse_iters, coarse_off;  // coarse_off: LDS bytes of the TC boxes (mode 1)
};

static_assert(sizeof(BotArgs<double, double>) <= 4096 && sizeof(BotArgs<float, double>) <= 4096,
              "bottom launch arguments exceed the kernel argument segment");

__device__ __forceinline__ int box_cells(const BotGeom &g) { return (g.nx + 2) * (g.ny + 2) * (g.nz + 2); }

// face ghosts of the cell at LDS index id = (i,j,k) with value v: beside it
// across a Neumann/Dirichlet face, or at the other end of a periodic axis
// (with an even number of cells there, the two ends are of different
// colours, so no thread of the colour being relaxed reads that ghost)
template <typename T>
__device__ __forceinline__ void face_ghosts(T *u, const BotGeom &g, const GhostL1<T> &gw, const BotPer &per, int id,
                                            int i, int j, int k, T v) {
  if (i == 0) {
    if (per.a[0]) u[id + g.nx * g.lxs] = v;
    else u[id - g.lxs] = gw.xf(1, v);
  }
  if (i == g.nx - 1) {
    if (per.a[0]) u[id - g.nx * g.lxs] = v;
    else u[id + g.lxs] = gw.xf(0, v);
  }
  if (j == 0) {
    if (per.a[1]) u[id + g.ny * g.lys] = v;
    else u[id - g.lys] = gw.xf(3, v);
  }
  if (j == g.ny - 1) {
    if (per.a[1]) u[id - g.ny * g.lys] = v;
    else u[id + g.lys] = gw.xf(2, v);
  }
  if (k == 0) {
    if (per.a[2]) u[id + g.nz] = v;
    else u[id - 1] = gw.xf(5, v);
  }
  if (k == g.nz - 1) {
    if (per.a[2]) u[id - g.nz] = v;
    else u[id + 1] = gw.xf(4, v);
  }
}

// A cell of a colour's compacted numbering t (threads over the colour's
// cells only): its LDS index, its face bits (x high, x low, y high, y low,
// z high, z low; bit 6: no cell) and its diagonal.  HOST: the
// relax_on_host numbering (every k, rows of ceil(nz/2) slots) and diagonal;
// else the device one (k < 2*(nz/2)).
template <typename D>
struct BotCell {
  int id, fl;
  D d;
};
constexpr int BOT_NONE = 64;

template <bool HOST, typename D, typename C>
__device__ __forceinline__ BotCell<D> bot_cell(const BotGeom &g, const C &c, int color, int t) {
  Mag mh;
  if constexpr (HOST) mh = g.mhh;
  else mh = g.mhd;
  const int h = (int)mh.n;
  const int row = mdiv(t, mh), i = mdiv(row, g.my), j = row - i * g.ny;
  const int k = 2 * (t - row * h) + ((i + j + color) & 1);
  BotCell<D> s;
  s.id = g.lshift + i * g.lxs + j * g.lys + k;
  s.fl = (i == g.nx - 1) | (i == 0) << 1 | (j == g.ny - 1) << 2 | (j == 0) << 3 | (k == g.nz - 1) << 4 |
         (k == 0) << 5 | (t >= g.nx * g.ny * h || k >= g.nz ? BOT_NONE : 0);
  if constexpr (HOST) {
    typedef decltype(c.omega) T;
    const T xm_ = (i == 0) ? c.m[1] : (i == g.nx - 1) ? c.m[0] : (T)0;
    const T ym_ = (j == 0) ? c.m[3] : (j == g.ny - 1) ? c.m[2] : (T)0;
    const T zm_ = (k == 0) ? c.m[5] : (k == g.nz - 1) ? c.m[4] : (T)0;
    s.d = diag_host(c, xm_, ym_, zm_);
  } else {
    s.d = diag_dev<D>(c, i, j, k, g.nx, g.ny, g.nz);
  }
  return s;
}

// face ghosts of a cell from its face bits (face_ghosts' writes): values
// and targets are computed for every face, so the transform coefficients
// are read before the branches, which only store
template <typename T>
__device__ __forceinline__ void face_ghosts_fl(T *u, const BotGeom &g, const GhostL1<T> &gw, const BotPer &per,
                                               int id, int fl, T v) {
  const int st[3] = {g.lxs, g.lys, 1}, nn[3] = {g.nx, g.ny, g.nz};
  T lv[3], hv[3];
  int lt[3], ht[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    lv[a] = per.a[a] ? v : gw.xf(2 * a + 1, v);
    hv[a] = per.a[a] ? v : gw.xf(2 * a, v);
    lt[a] = per.a[a] ? id + nn[a] * st[a] : id - st[a];
    ht[a] = per.a[a] ? id - nn[a] * st[a] : id + st[a];
  }
#pragma unroll
  for (int a = 0; a < 3; a++) {
    if (fl & (2 << (2 * a))) u[lt[a]] = lv[a];
    if (fl & (1 << (2 * a))) u[ht[a]] = hv[a];
  }
}

// relax one cell (device or relax_on_host arithmetic) and its face ghosts
template <bool HOST, typename T, typename D, typename C>
__device__ __forceinline__ void bot_relax_cell(T *u, const T *b, const BotGeom &g, const C &c, const GhostL1<T> &gw,
                                               const BotPer &per, const BotCell<D> &s) {
  const int id = s.id;
  T v;
  if constexpr (HOST)
    v = relax_pt_host_d(c, u[id], u[id - 1], u[id + 1], u[id - g.lys], u[id + g.lys], u[id - g.lxs], u[id + g.lxs],
                        b[id], s.d);
  else
    v = relax_pt(c, u[id], u[id - 1], u[id + 1], u[id - g.lys], u[id + g.lys], u[id - g.lxs], u[id + g.lxs], b[id],
                 s.d);
  u[id] = v;
  if (s.fl) face_ghosts_fl(u, g, gw, per, id, s.fl, v);
}

// iters red-black sweeps of a level (HOST: the coarsest's relax_on_host
// arithmetic).  Up to BOT_SLOTS cells of each colour per thread are decoded
// once and kept in registers for all sweeps; larger levels decode per
// sweep.  The face ghosts follow every update, so the 7-point reads of the
// level need no fill (edges and corners are left to the fill before
// anything reads them).
constexpr int BOT_SLOTS = 2;
template <bool HOST, typename T, typename C>
__device__ __forceinline__ void bot_sweeps(T *u, const T *b, const BotGeom &g, const C &c, const GhostL1<T> &gw_arg,
                                           const BotPer &per, int iters) {
  GhostL1<T> gw;
#pragma unroll
  for (int f = 0; f < 6; f++) gw.s[f] = held(gw_arg.s[f]), gw.add[f] = held(gw_arg.add[f]);
  int n;
  if constexpr (HOST) n = g.nx * g.ny * (int)g.mhh.n;
  else n = g.nx * g.ny * (int)g.mhd.n;
  if (n <= BOT_SLOTS * (int)blockDim.x) {
    BotCell<T> cell[2][BOT_SLOTS];
#pragma unroll
    for (int col = 0; col < 2; col++)
#pragma unroll
      for (int q = 0; q < BOT_SLOTS; q++) cell[col][q] = bot_cell<HOST, T>(g, c, col, threadIdx.x + q * blockDim.x);
    for (int it = 0; it < iters; it++) {
#pragma unroll
      for (int col = 0; col < 2; col++) {
#pragma unroll
        for (int q = 0; q < BOT_SLOTS; q++)
          if (!(cell[col][q].fl & BOT_NONE)) bot_relax_cell<HOST>(u, b, g, c, gw, per, cell[col][q]);
        __syncthreads();
      }
    }
  } else {
    for (int it = 0; it < iters; it++)
      for (int col = 0; col < 2; col++) {
        for (int t = threadIdx.x; t < n; t += blockDim.x) {
          const BotCell<T> s = bot_cell<HOST, T>(g, c, col, t);
          if (!(s.fl & BOT_NONE)) bot_relax_cell<HOST>(u, b, g, c, gw, per, s);
        }
        __syncthreads();
      }
  }
}

// the whole level-1 ghost layer (faces, edges, corners) from the interior:
// out-of-range indices clamp to the adjacent cell and take that face's
// transform, x then y then z, or on a periodic axis wrap to the other end
// untransformed (the reference fill's x, y, z passes)
template <typename T>
__device__ void bot_fill(T *u, const BotGeom &g, const GhostL1<T> &gw, const BotPer &per) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int nZ = (nx + 2) * (ny + 2), nY = (nx + 2) * nz, nX = ny * nz;
  for (int t = threadIdx.x; t < 2 * (nZ + nY + nX); t += blockDim.x) {
    int i, j, k, q = t;
    if (q < 2 * nZ) {
      const int side = q >= nZ;
      q -= side * nZ;
      i = mdiv(q, g.mpy);
      j = q - i * (ny + 2) - 1;
      i -= 1;
      k = side ? nz : -1;
    } else if ((q -= 2 * nZ) < 2 * nY) {
      const int side = q >= nY;
      q -= side * nY;
      i = mdiv(q, g.mz);
      k = q - i * nz;
      i -= 1;
      j = side ? ny : -1;
    } else {
      q -= 2 * nY;
      const int side = q >= nX;
      q -= side * nX;
      j = mdiv(q, g.mz);
      k = q - j * nz;
      i = side ? nx : -1;
    }
    const int sx = i < 0 ? -1 : i >= nx ? 1 : 0, sy = j < 0 ? -1 : j >= ny ? 1 : 0, sz = k < 0 ? -1 : k >= nz ? 1 : 0;
    const int si = per.a[0] ? i - sx * nx : i - sx, sj = per.a[1] ? j - sy * ny : j - sy,
              sk = per.a[2] ? k - sz * nz : k - sz;
    const int src = g.lshift + si * g.lxs + sj * g.lys + sk;
    u[g.lshift + i * g.lxs + j * g.lys + k] =
        gw.chain(per.a[0] ? 0 : sx, per.a[1] ? 0 : sy, per.a[2] ? 0 : sz, u[src]);
  }
}

template <typename T>
__device__ __forceinline__ void bot_relax(T *u, const T *b, const BotLevel<T> &L, const BotPer &per, int iters) {
  bot_sweeps<false>(u, b, L.g, L.dc, L.gw, per, iters);
}

// b_c = restrict(b_f - A u_f) over the coarse interior (k_res_restrict)
template <typename T>
__device__ void bot_res_restrict(const T *uf, const T *bf, const BotGeom &gf, const ResC<T> &c, T *bc,
                                 const BotGeom &gc) {
  for (int t = threadIdx.x; t < (int)gc.mn.n; t += blockDim.x) {
    const int row = mdiv(t, gc.mz), i = mdiv(row, gc.my), j = row - i * gc.ny, k = t - row * gc.nz;
    T rr[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int id = gf.lshift + (2 * i + (q >> 2)) * gf.lxs + (2 * j + ((q >> 1) & 1)) * gf.lys + 2 * k + (q & 1);
      rr[q] = bf[id] - (c.fz * (uf[id - 1] + uf[id + 1]) + c.fy * (uf[id - gf.lys] + uf[id + gf.lys]) +
                        c.fx * (uf[id - gf.lxs] + uf[id + gf.lxs]) - c.diag * uf[id]);
    }
    bc[gc.lshift + i * gc.lxs + j * gc.lys + k] = restrict_sum(rr[0], rr[1], rr[2], rr[3], rr[4], rr[5], rr[6], rr[7]);
  }
}

// u_f += trilinear prolongation of u_c (k_prolong, reference corner choice);
// the face ghosts of every updated cell follow it (what the level-1 fill
// after the prolongation would give them)
template <typename T>
__device__ void bot_prolong(const T *uc, const BotGeom &gc, T *uf, const BotGeom &gf, const GhostL1<T> &gw,
                            const BotPer &per) {
  const int n = (int)gc.mn.n;
  for (int t = threadIdx.x; t < 8 * n; t += blockDim.x) {
    const int q = mdiv(t, gc.mn), ct = t - q * n;  // child-major: consecutive threads walk coarse cells
    const int row = mdiv(ct, gc.mz), ic = mdiv(row, gc.my), jc = row - ic * gc.ny, kc = ct - row * gc.nz;
    const int dx = q >> 2, dy = (q >> 1) & 1, dz = q & 1;
    const int id = gc.lshift + ic * gc.lxs + jc * gc.lys + kc;
    const int sx = dx ? gc.lxs : -gc.lxs, sy = dy ? gc.lys : -gc.lys, sz = dz ? 1 : -1;
    const int cx = (dx == 0 && dy == 1 && dz == 1) ? gc.lxs : sx;
    const T C = uc[id];
    const T X = uc[id + sx], Y = uc[id + sy], Z = uc[id + sz];
    const T XY = uc[id + sx + sy], XZ = uc[id + sx + sz], YZ = uc[id + sy + sz];
    const T XYZ = uc[id + cx + sy + sz];
    const int i = 2 * ic + dx, j = 2 * jc + dy, k = 2 * kc + dz;
    const int idf = gf.lshift + i * gf.lxs + j * gf.lys + k;
    const T v = uf[idf] + PW<T>::c0 * C + PW<T>::c1 * (X + Y + Z) + PW<T>::c2 * (XY + XZ + YZ) + PW<T>::c3 * XYZ;
    uf[idf] = v;
    face_ghosts(uf, gf, gw, per, idf, i, j, k, v);
  }
}

// Global boxes (ghosts included) -> LDS: NG grids of the same geometry
// (u and b of the top level, or b alone), BOT_LOADS cells per thread and
// grid at a time, every global load in flight before the first LDS store
// waits for one.  And LDS box -> global grid.
constexpr int BOT_LOADS = 8;
template <typename T>
__device__ __forceinline__ T &box_at(const DGrid<T> &gr, const BotGeom &g, int t) {
  const int row = mdiv(t, g.mpz), i = mdiv(row, g.mpy), j = row - i * (g.ny + 2), k = t - row * (g.nz + 2);
  return gr.at(i - 1, j - 1, k - 1);
}
template <int NG, typename T>
__device__ void bot_load(T *const (&s)[NG], const DGrid<T> (&gr)[NG], const BotGeom &g) {
  const int n = box_cells(g);
  for (int t0 = threadIdx.x; t0 < n; t0 += BOT_LOADS * blockDim.x) {
    T v[NG][BOT_LOADS];
#pragma unroll
    for (int q = 0; q < BOT_LOADS; q++) {
      const int t = t0 + q * blockDim.x;
      const int tc = min(t, n - 1);  // past the box: a redundant load, no store
#pragma unroll
      for (int a = 0; a < NG; a++) v[a][q] = box_at(gr[a], g, tc);
    }
#pragma unroll
    for (int q = 0; q < BOT_LOADS; q++) {
      const int t = t0 + q * blockDim.x;
#pragma unroll
      for (int a = 0; a < NG; a++)
        if (t < n) s[a][t] = v[a][q];
    }
  }
}
template <typename T>
__device__ void bot_store_box(const T *s, const DGrid<T> &gr, const BotGeom &g) {
  for (int t = threadIdx.x; t < box_cells(g); t += blockDim.x) box_at(gr, g, t) = s[t];
}
template <typename T>
__device__ void bot_store_interior(const T *s, const DGrid<T> &gr, const BotGeom &g) {
  for (int t = threadIdx.x; t < (int)g.mn.n; t += blockDim.x) {
    const int row = mdiv(t, g.mz), i = mdiv(row, g.my), j = row - i * g.ny, k = t - row * g.nz;
    gr.at(i, j, k) = s[g.lshift + i * g.lxs + j * g.lys + k];
  }
}
template <typename T>
__device__ void bot_zero(T *s, int n) {
  for (int t = threadIdx.x; t < n; t += blockDim.x) s[t] = (T)0;
}

template <typename T, typename TC, int MODE>
__global__ __launch_bounds__(BOT_THREADS) void k_mg_bottom(BotArgs<T, TC> args) {
  // the arguments are read where they lie, in the kernel-argument segment
  // (the by-value parameter is its first and only explicit argument): a
  // private copy of a struct this large ends up in scratch memory
  (void)args;
  const BotArgs<T, TC> &A = *(const BotArgs<T, TC> *)__builtin_amdgcn_kernarg_segment_ptr();
  extern __shared__ unsigned char smem[];
  auto U = [&](int l) { return (T *)(smem + A.L[l].off); };
  auto B = [&](int l) { return (T *)(smem + A.L[l].off) + box_cells(A.L[l].g); };
  const int C = A.nlev - 1;

  // top level: u (or the zero grid with its zero ghosts) and b; the u and
  // b boxes below start at zero (the b ghosts stay zero)
  if (A.zero_top) {
    bot_zero(U(0), box_cells(A.L[0].g));
    T *const s[1] = {B(0)};
    const DGrid<T> gr[1] = {A.L[0].b};
    bot_load<1>(s, gr, A.L[0].g);
  } else {
    T *const s[2] = {U(0), B(0)};
    const DGrid<T> gr[2] = {A.L[0].u, A.L[0].b};
    bot_load<2>(s, gr, A.L[0].g);
  }
  for (int l = 1; l <= C; l++) bot_zero(U(l), 2 * box_cells(A.L[l].g));
  __syncthreads();

  // down: relax, residual -> restriction (the residual reads face ghosts
  // only, which the relaxation keeps; without a sweep they are refilled)
  for (int l = 0; l < C; l++) {
    bot_relax(U(l), B(l), A.L[l], A.per, A.nu1);
    if (A.nu1 == 0) {
      bot_fill(U(l), A.L[l].g, A.L[l].gw, A.per);
      __syncthreads();
    }
    bot_res_restrict(U(l), B(l), A.L[l].g, A.L[l].rc, B(l + 1), A.L[l + 1].g);
    __syncthreads();
  }

  // coarsest level: relax_on_host iterations on the zero start, then the
  // level-1 fill (mode 1: in TC on copies of the boxes)
  {
    const BotGeom &g = A.L[C].g;
    const int box = box_cells(g);
    TC *uc = (TC *)U(C), *bc = (TC *)B(C);
    if (MODE == 1) {
      uc = (TC *)(smem + A.coarse_off);
      bc = uc + box;
      for (int t = threadIdx.x; t < box; t += blockDim.x) {
        uc[t] = (TC)U(C)[t];
        bc[t] = (TC)B(C)[t];
      }
      __syncthreads();
    }
    bot_sweeps<true>(uc, bc, g, A.hc, A.gwc, A.per, A.coarse_iters);
    bot_fill(uc, g, A.gwc, A.per);
    __syncthreads();
    if (MODE == 1) {
      for (int t = threadIdx.x; t < box; t += blockDim.x) U(C)[t] = (T)uc[t];
      __syncthreads();
    }
  }

  // up: prolongation (with its face ghosts), relax, level-1 fill (the
  // next prolongation reads edge and corner ghosts of this level)
  for (int l = C - 1; l >= 0; l--) {
    bot_prolong(U(l + 1), A.L[l + 1].g, U(l), A.L[l].g, A.L[l].gw, A.per);
    __syncthreads();
    bot_relax(U(l), B(l), A.L[l], A.per, A.nu2);
    bot_fill(U(l), A.L[l].g, A.L[l].gw, A.per);
    __syncthreads();
  }

  // write back every level (u boxes, b interiors below the top)
  for (int l = 0; l <= C; l++) {
    bot_store_box(U(l), A.L[l].u, A.L[l].g);
    if (l > 0) bot_store_interior(B(l), A.L[l].b, A.L[l].g);
  }
}

FacesS faces_s(const ocu_bc_set *bc) {
  FacesS F;
  const ocu_bc *all[6] = {&bc->xpos, &bc->xneg, &bc->ypos, &bc->yneg, &bc->zpos, &bc->zneg};
  for (int q = 0; q < 6; q++) F.f[q] = {all[q]->type, all[q]->value};
  return F;
}

int check_pair(const ocu_grid3d *u, const ocu_grid3d *b, const char *what) {
  if (int e = check_grid(u, what)) return e;
  if (int e = check_grid(b, what)) return e;
  OCU_REQUIRE(u->dtype == b->dtype && u->dtype != OCU_I32, "%s: dtype mismatch", what);
  OCU_REQUIRE(u->nx == b->nx && u->ny == b->ny && u->nz == b->nz, "%s: interior dimension mismatch", what);
  OCU_REQUIRE(u->gx >= 1 && u->gy >= 1 && u->gz >= 1, "%s: needs one ghost layer", what);
  return OCU_OK;
}

// Lanes per cell of the whole-row fused kernel for this problem, or 0 when
// the tiled kernel must run: the row must fit one wave (nz <= 256), split
// evenly over the lanes, and every V-vector access must be aligned.
int row_vec_width(const ocu_grid3d *u, const ocu_grid3d *o, const ocu_grid3d *b) {
  const int nz = u->nz;
  const int V = nz <= 64 ? 1 : nz <= 128 ? 2 : nz <= 256 ? 4 : 0;
  if (!V || nz % V) return 0;
  const size_t w = u->dtype == OCU_F64 ? 8 : 4;
  for (const ocu_grid3d *g : {u, o, b}) {
    if (g->pnz % V || ((int64_t)g->pny * g->pnz) % V) return 0;
    if (((uintptr_t)g->base + (uintptr_t)g->shift * w) % (w * V)) return 0;
  }
  return V;
}

// Resident blocks of a kernel on the current device: CUs and blocks per CU,
// cached per device, kernel and block size.
struct Slots {
  int cus, per;
};
template <typename K>
Slots device_slots(K kern, int threads) {
  struct Entry {
    const void *k;
    int dev, threads;
    Slots s;
  };
  static std::mutex mu;
  static Entry cache[128];
  static int used = 0;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0) return Slots{256, 1};
  const void *key = reinterpret_cast<const void *>(kern);
  std::lock_guard<std::mutex> lock(mu);
  for (int q = 0; q < used; q++)
    if (cache[q].k == key && cache[q].dev == dev && cache[q].threads == threads) return cache[q].s;
  int cus = 0, per = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, threads, 0) != hipSuccess || per <= 0) per = 1;
  const Slots s{cus, per};
  if (used < 128) cache[used++] = Entry{key, dev, threads, s};
  return s;
}

// i-planes per block of the whole-row kernel.  A block reads ichunk+3
// planes to write ichunk (the two-colour halo).  Two costs compete: the
// march of one block (rounds of resident blocks x planes per block, latency
// bound) and the planes every CU streams in total (nblocks x planes / CUs,
// bandwidth bound); the chunk minimising 3 x march + 2 x stream is taken,
// weights fitted to 32^3..256^3 sweeps.  The OCU_FUSED_ICHUNK environment
// variable overrides the choice (tuning).
int fused_ichunk(int nx, int tiles_j, Slots sl) {
  if (const char *s = getenv("OCU_FUSED_ICHUNK")) {
    int v = atoi(s);
    if (v > 0) return v;
  }
  static const int cand[] = {1, 2, 4, 6, 8, 12, 16, 24, 32, 48, 64};
  const int64_t slots = (int64_t)sl.cus * sl.per;
  int best = 1;
  int64_t best_cost = INT64_MAX;
  for (int ic : cand) {
    const int64_t nb = (int64_t)tiles_j * ((nx + ic - 1) / ic);
    const int64_t planes = min(ic, nx) + 3;
    const int64_t cost = 3 * ((nb + slots - 1) / slots) * planes * sl.cus + 2 * nb * planes;
    if (cost < best_cost) best_cost = cost, best = ic;
    if (ic >= nx) break;
  }
  return best;
}

template <typename T, int V, bool ZERO, bool PRO = false>
int launch_fused_rows(const ocu_grid3d *u_in, const ocu_grid3d *u_out, const ocu_grid3d *b,
                      const typename DevCFor<T>::type &c, int first_color, const GhostL1<T> &gw, int wghost,
                      hipStream_t stream, const ocu_grid3d *uc = nullptr, int quirk = 0, int perx = 0,
                      int pery = 0) {
  typedef RowsShape<T, V, PRO> RS;
  constexpr int NR = RS::R * RS::NW;
  // the periodic variant only where a face pair is periodic: the plain
  // sweep carries none of its wrap logic
  auto kern = (!PRO && (perx || pery)) ? k_relax_fused_rows<T, V, RS::R, RS::NW, ZERO, PRO, !PRO>
                                       : k_relax_fused_rows<T, V, RS::R, RS::NW, ZERO, PRO, false>;
  const int nx = u_in->nx, ny = u_in->ny;
  const int tiles_j = (ny + NR - 5) / (NR - 4);
  const int ichunk = fused_ichunk(nx, tiles_j, device_slots(kern, 64 * RS::NW));
  const int ntiles = tiles_j * ((nx + ichunk - 1) / ichunk);
  const int nblocks = ((ntiles + FXCD - 1) / FXCD) * FXCD;
  kern<<<nblocks, dim3(64, RS::NW), 0, stream>>>(dgrid<T>(u_in), dgrid<T>(u_out), dgrid<T>(b), c, first_color,
                                                 tiles_j, ichunk, ntiles, gw, wghost,
                                                 PRO ? dgrid<T>(uc) : dgrid<T>(u_in), quirk, perx, pery);
  OCU_HIP_LAUNCHED("ocu_mg_relax_fused");
}

template <typename T, bool ZERO>
int launch_fused_tiled(const ocu_grid3d *u_in, const ocu_grid3d *u_out, const ocu_grid3d *b,
                       const typename DevCFor<T>::type &c, int first_color, const GhostL1<T> &gw, int wghost,
                       hipStream_t stream) {
  const int nx = u_in->nx, ny = u_in->ny, nz = u_in->nz;
  const int tiles_k = (nz + FOK - 1) / FOK, tiles_j = (ny + FOJ - 1) / FOJ;
  // enough i-chunks to give every CU a few blocks, but long enough marches
  const int ichunk = nx >= 256 ? 16 : nx >= 64 ? 8 : 4;
  const int ntiles = tiles_k * tiles_j * ((nx + ichunk - 1) / ichunk);
  const int nblocks = ((ntiles + FXCD - 1) / FXCD) * FXCD;
  k_relax_fused<T, ZERO><<<nblocks, dim3(FBK, FBJ), 0, stream>>>(dgrid<T>(u_in), dgrid<T>(u_out), dgrid<T>(b), c,
                                                                 first_color, tiles_k, tiles_j, ichunk, ntiles, gw,
                                                                 wghost);
  OCU_HIP_LAUNCHED("ocu_mg_relax_fused");
}

template <typename T, bool ZERO>
int launch_fused(const ocu_grid3d *u_in, const ocu_grid3d *u_out, const ocu_grid3d *b,
                 const typename DevCFor<T>::type &c, int first_color, const GhostL1<T> &gw, int wghost,
                 hipStream_t stream, int perx = 0, int pery = 0) {
  switch (row_vec_width(u_in, u_out, b)) {
    case 1:
      return launch_fused_rows<T, 1, ZERO>(u_in, u_out, b, c, first_color, gw, wghost, stream, nullptr, 0, perx, pery);
    case 2:
      return launch_fused_rows<T, 2, ZERO>(u_in, u_out, b, c, first_color, gw, wghost, stream, nullptr, 0, perx, pery);
    case 4:
      return launch_fused_rows<T, 4, ZERO>(u_in, u_out, b, c, first_color, gw, wghost, stream, nullptr, 0, perx, pery);
    default: return launch_fused_tiled<T, ZERO>(u_in, u_out, b, c, first_color, gw, wghost, stream);
  }
}

// The whole-row fused sweep also takes periodic x and/or y face pairs (even
// counts along them) with Neumann/Dirichlet z faces, without the zero-input
// mode: 1 when that applies to this grid and these faces.
int fused_periodic_ok(const ocu_grid3d *u, const ocu_grid3d *o, const ocu_grid3d *b, const ocu_bc_set *bc) {
  if (!gfaces_paired(bc) || bc->zpos.type == OCU_BC_PERIODIC) return 0;
  const int px = bc->xpos.type == OCU_BC_PERIODIC, py = bc->ypos.type == OCU_BC_PERIODIC;
  if (!px && !py) return 0;
  if ((px && u->nx % 2) || (py && u->ny % 2)) return 0;
  return row_vec_width(u, o, b) ? 1 : 0;
}

int relax_fused(const ocu_grid3d *u_in, const ocu_grid3d *u_out, const ocu_grid3d *b, const ocu_mg_level_params *p,
                const ocu_bc_set *bc, double hx, double hy, double hz, int first_color, int flags, void *stream) {
  const char *what = "ocu_mg_relax_fused";
  if (int e = check_pair(u_in, b, what)) return e;
  if (int e = check_grid(u_out, "ocu_mg_relax_fused(out)")) return e;
  OCU_REQUIRE(layout_match(u_in, u_out) && u_in->dtype == u_out->dtype && u_in->nx == u_out->nx &&
                  u_in->ny == u_out->ny && u_in->nz == u_out->nz,
              "%s: u_out must be congruent with u_in", what);
  OCU_REQUIRE(u_out->gx >= 1 && u_out->gy >= 1 && u_out->gz >= 1, "%s: u_out needs one ghost layer", what);
  OCU_REQUIRE((flags & ~(OCU_FUSED_WRITE_GHOSTS | OCU_FUSED_ZERO_INPUT)) == 0, "%s: bad flags %d", what, flags);
  const bool zero = flags & OCU_FUSED_ZERO_INPUT;
  const int wghost = (flags & OCU_FUSED_WRITE_GHOSTS) ? 1 : 0;
  OCU_REQUIRE(zero || u_in->base != u_out->base, "%s: u_out must not alias u_in", what);
  OCU_REQUIRE(p && bc, "%s: null params", what);
  OCU_REQUIRE(first_color == 0 || first_color == 1, "%s: bad first_color %d", what, first_color);
  const bool periodic = any_periodic(bc);
  if (periodic && !fused_periodic_ok(u_in, u_out, b, bc)) {
    set_error("%s: periodic faces need x/y pairs with even counts, non-periodic z and whole rows", what);
    return OCU_ERR_UNSUPPORTED;
  }
  if (!periodic && (wghost || zero) && !gfaces_local(bc)) {
    set_error("%s: ghost writes and zero input need Neumann or Dirichlet faces", what);
    return OCU_ERR_UNSUPPORTED;
  }
  if (zero && !gfaces_homogeneous(bc)) {
    set_error("%s: zero input needs homogeneous faces (value 0)", what);
    return OCU_ERR_UNSUPPORTED;
  }
  if ((int64_t)u_in->nx * u_in->ny * u_in->nz == 0) return OCU_OK;
  const int px = bc->xpos.type == OCU_BC_PERIODIC, py = bc->ypos.type == OCU_BC_PERIODIC;
  if (u_in->dtype == OCU_F64) {
    const DevC64 c = dev_c64(p);
    const GhostL1<double> gw = ghost_l1<double>(bc, hx, hy, hz);
    return zero ? launch_fused<double, true>(u_in, u_out, b, c, first_color, gw, wghost, S(stream), px, py)
                : launch_fused<double, false>(u_in, u_out, b, c, first_color, gw, wghost, S(stream), px, py);
  }
  const DevC32 c = dev_c32(p);
  const GhostL1<float> gw = ghost_l1<float>(bc, hx, hy, hz);
  return zero ? launch_fused<float, true>(u_in, u_out, b, c, first_color, gw, wghost, S(stream), px, py)
              : launch_fused<float, false>(u_in, u_out, b, c, first_color, gw, wghost, S(stream), px, py);
}

// The whole-row width of the prolongation-fused sweep (4), or 0 when it
// cannot run: fine exactly twice the coarse grid, rows of four cells per
// lane (129..256 cells; on shorter rows the separate prolongation is the
// faster of the two), coarse rows aligned for half-width vectors, faces that
// the ghost transforms cover, and a coarse ghost layer.
int fused_prolong_width(const ocu_grid3d *u_in, const ocu_grid3d *u_out, const ocu_grid3d *b,
                        const ocu_grid3d *uc, const ocu_bc_set *bc) {
  if (!uc || !bc || any_periodic(bc) || !gfaces_local(bc)) return 0;
  if (uc->dtype != u_in->dtype || u_in->nx != 2 * uc->nx || u_in->ny != 2 * uc->ny || u_in->nz != 2 * uc->nz)
    return 0;
  if (uc->gx < 1 || uc->gy < 1 || uc->gz < 1) return 0;
  const int V = row_vec_width(u_in, u_out, b);
  if (V != 4 || !row_aligned(uc, V / 2)) return 0;
  return V;
}

// fine cells per lane of k_res_restrict_rows (whole rows, whole coarse
// cells per lane, aligned vectors), or 0 when the tiled kernel must run
int res_row_width(const ocu_grid3d *u, const ocu_grid3d *b, const ocu_grid3d *bc) {
  const int nz = u->nz;
  const int V = nz <= 128 ? 2 : nz <= 256 ? 4 : 0;
  if (!V || nz % V || u->gx < 1 || u->gy < 1 || u->gz < 1) return 0;
  const size_t w = u->dtype == OCU_F64 ? 8 : 4;
  auto aligned = [&](const ocu_grid3d *g, int v) {
    return g->pnz % v == 0 && ((int64_t)g->pny * g->pnz) % v == 0 &&
           ((uintptr_t)g->base + (uintptr_t)g->shift * w) % (w * v) == 0;
  };
  return aligned(u, V) && aligned(b, V) && aligned(bc, V / 2) ? V : 0;
}

template <int V, typename T>
int launch_res_restrict_rows_t(const ocu_grid3d *u, const ocu_grid3d *b, const ocu_grid3d *bc, double h, double fx,
                               double fy, double fz, double *norm_ws, hipStream_t stream) {
  constexpr int NW = 8;
  const int tiles_j = (bc->ny + NW - 1) / NW;
  // about 4096 waves in all: enough to fill the chip, long marches; with
  // norm partials at most NORM_PARTS tiles (one partial pair per tile)
  int ichunk = std::max(1, std::min(bc->nx, (int)(((int64_t)bc->nx * bc->ny) / 4096)));
  if (norm_ws) ichunk = std::max(ichunk, (int)(((int64_t)tiles_j * bc->nx + NORM_PARTS - 1) / NORM_PARTS));
  const int ntiles = tiles_j * ((bc->nx + ichunk - 1) / ichunk);
  const ResC<T> c = res_c<T>(h, fx, fy, fz);
  if (norm_ws) {
    // NORM_PARTS is a multiple of FXCD and ntiles <= NORM_PARTS
    k_res_restrict_rows<T, V, NW, true><<<NORM_PARTS, dim3(64, NW), 0, stream>>>(
        dgrid<T>(u), dgrid<T>(b), dgrid<T>(bc), c, tiles_j, ichunk, ntiles, norm_ws);
  } else {
    const int nblocks = ((ntiles + FXCD - 1) / FXCD) * FXCD;
    k_res_restrict_rows<T, V, NW, false><<<nblocks, dim3(64, NW), 0, stream>>>(
        dgrid<T>(u), dgrid<T>(b), dgrid<T>(bc), c, tiles_j, ichunk, ntiles, nullptr);
  }
  OCU_HIP_LAUNCHED("ocu_mg_residual_restrict");
}
template <int V>
int launch_res_restrict_rows(const ocu_grid3d *u, const ocu_grid3d *b, const ocu_grid3d *bc, double h, double fx,
                             double fy, double fz, double *norm_ws, hipStream_t stream) {
  return u->dtype == OCU_F64 ? launch_res_restrict_rows_t<V, double>(u, b, bc, h, fx, fy, fz, norm_ws, stream)
                             : launch_res_restrict_rows_t<V, float>(u, b, bc, h, fx, fy, fz, norm_ws, stream);
}

template <typename T, int V>
int launch_prolong_rows(const ocu_grid3d *uc, const ocu_grid3d *uf, int quirk, const GhostL1<T> &gw, int wghost,
                        hipStream_t stream, int perx = 0, int pery = 0) {
  constexpr int NW = 8;
  const int tiles_j = (uf->ny + NW - 1) / NW;
  // coarse planes per wave: about 4096 waves in all
  const int ichunk = std::max(1, std::min(uc->nx, (int)(((int64_t)uc->nx * uf->ny) / 4096)));
  const int ntiles = tiles_j * ((uc->nx + ichunk - 1) / ichunk);
  const int nblocks = ((ntiles + FXCD - 1) / FXCD) * FXCD;
  k_prolong_rows<T, V, NW><<<nblocks, dim3(64, NW), 0, stream>>>(dgrid<T>(uc), dgrid<T>(uf), quirk, gw, wghost,
                                                                 tiles_j, ichunk, ntiles, perx, pery);
  OCU_HIP_LAUNCHED("ocu_mg_prolong");
}

// ghost writes of the prolongation: Neumann/Dirichlet faces (either
// kernel), or periodic x/y pairs with non-periodic z on the whole-row kernel
// (fine exactly twice the coarse grid, so the periodic counts are even)
bool prolong_bc_ok(const ocu_grid3d *u_coarse, const ocu_grid3d *u_fine, const ocu_bc_set *bc) {
  if (!any_periodic(bc)) return gfaces_local(bc);
  if (!gfaces_paired(bc) || bc->zpos.type == OCU_BC_PERIODIC) return false;
  if (u_fine->nx != 2 * u_coarse->nx || u_fine->ny != 2 * u_coarse->ny || u_fine->nz != 2 * u_coarse->nz) return false;
  return res_row_width(u_fine, u_fine, u_coarse) != 0;
}

int prolong(const ocu_grid3d *u_coarse, const ocu_grid3d *u_fine, int quirk, const ocu_bc_set *bc, double hx,
            double hy, double hz, void *stream) {
  if (int e = check_grid(u_coarse, "ocu_mg_prolong(coarse)")) return e;
  if (int e = check_grid(u_fine, "ocu_mg_prolong(fine)")) return e;
  OCU_REQUIRE(u_coarse->dtype == u_fine->dtype && u_fine->dtype != OCU_I32, "ocu_mg_prolong: dtype mismatch");
  OCU_REQUIRE(u_fine->nx >= 2 * u_coarse->nx && u_fine->ny >= 2 * u_coarse->ny && u_fine->nz >= 2 * u_coarse->nz,
              "ocu_mg_prolong: fine grid must be twice the coarse grid");
  OCU_REQUIRE(u_coarse->gx >= 1 && u_coarse->gy >= 1 && u_coarse->gz >= 1, "ocu_mg_prolong: coarse needs ghosts");
  const int wghost = bc ? 1 : 0;
  int perx = 0, pery = 0;
  if (wghost) {
    // every fine boundary cell must be produced by this launch
    OCU_REQUIRE(u_fine->nx == 2 * u_coarse->nx && u_fine->ny == 2 * u_coarse->ny && u_fine->nz == 2 * u_coarse->nz,
                "ocu_mg_prolong_bc: the fine grid must be exactly twice the coarse grid");
    OCU_REQUIRE(u_fine->gx >= 1 && u_fine->gy >= 1 && u_fine->gz >= 1, "ocu_mg_prolong_bc: fine needs ghosts");
    if (!prolong_bc_ok(u_coarse, u_fine, bc)) {
      set_error("ocu_mg_prolong_bc: ghost writes need Neumann or Dirichlet faces, or periodic x/y pairs on "
                "whole rows");
      return OCU_ERR_UNSUPPORTED;
    }
    perx = bc->xpos.type == OCU_BC_PERIODIC;
    pery = bc->ypos.type == OCU_BC_PERIODIC;
  }
  if ((int64_t)u_coarse->nx * u_coarse->ny * u_coarse->nz == 0) return OCU_OK;
  static const ocu_bc_set none = {};
  const ocu_bc_set *fb = bc ? bc : &none;
  if (u_fine->nx == 2 * u_coarse->nx && u_fine->ny == 2 * u_coarse->ny && u_fine->nz == 2 * u_coarse->nz) {
    const bool f64 = u_fine->dtype == OCU_F64;
    switch (res_row_width(u_fine, u_fine, u_coarse)) {
      case 2:
        return f64 ? launch_prolong_rows<double, 2>(u_coarse, u_fine, quirk, ghost_l1<double>(fb, hx, hy, hz), wghost,
                                                     S(stream), perx, pery)
                   : launch_prolong_rows<float, 2>(u_coarse, u_fine, quirk, ghost_l1<float>(fb, hx, hy, hz), wghost,
                                                    S(stream), perx, pery);
      case 4:
        return f64 ? launch_prolong_rows<double, 4>(u_coarse, u_fine, quirk, ghost_l1<double>(fb, hx, hy, hz), wghost,
                                                     S(stream), perx, pery)
                   : launch_prolong_rows<float, 4>(u_coarse, u_fine, quirk, ghost_l1<float>(fb, hx, hy, hz), wghost,
                                                    S(stream), perx, pery);
      default: break;
    }
  }
  dim3 grid = stencil_grid(u_coarse->nx, u_coarse->ny, u_coarse->nz);
  if (u_fine->dtype == OCU_F64)
    k_prolong<double><<<grid, stencil_block(), 0, S(stream)>>>(dgrid<double>(u_coarse), dgrid<double>(u_fine), quirk,
                                                               ghost_l1<double>(fb, hx, hy, hz), wghost);
  else
    k_prolong<float><<<grid, stencil_block(), 0, S(stream)>>>(dgrid<float>(u_coarse), dgrid<float>(u_fine), quirk,
                                                              ghost_l1<float>(fb, hx, hy, hz), wghost);
  OCU_HIP_LAUNCHED("ocu_mg_prolong");
}

// LDS layout of a bottom launch: per level the u box then the b box (T),
// then (mode 1) the TC copies of the coarsest boxes.  Returns the bytes.
template <typename T, typename TC>
size_t bottom_layout(const ocu_grid3d *u, int nlev, int mode, BotArgs<T, TC> *A) {
  size_t off = 0;
  for (int l = 0; l < nlev; l++) {
    const int nx = u[l].nx, ny = u[l].ny, nz = u[l].nz;
    if (A) {
      BotGeom &g = A->L[l].g;
      g.nx = nx; g.ny = ny; g.nz = nz;
      g.lys = nz + 2;
      g.lxs = (ny + 2) * g.lys;
      g.lshift = g.lxs + g.lys + 1;
      g.my = mag(ny); g.mz = mag(nz); g.mpy = mag(ny + 2); g.mpz = mag(nz + 2);
      g.mhd = mag(nz >> 1 ? nz >> 1 : 1);
      g.mhd.n = nz >> 1;  // slot count; the divisor is unused when zero
      g.mhh = mag((nz + 1) / 2);
      g.mn = mag(nx * ny * nz);
      A->L[l].off = (int)off;
    }
    off += 2 * (size_t)(nx + 2) * (ny + 2) * (nz + 2) * sizeof(T);
    off = (off + 15) & ~(size_t)15;
  }
  if (A) A->coarse_off = (int)off;
  if (mode == 1) {
    const ocu_grid3d &c = u[nlev - 1];
    off += 2 * (size_t)(c.nx + 2) * (c.ny + 2) * (c.nz + 2) * sizeof(TC);
  }
  return off;
}

// LDS of one workgroup on gfx950 (160 KiB per CU, all of it available to a
// single workgroup)
constexpr size_t BOT_LDS_MAX = 163840;

template <typename T, typename TC, int MODE>
int launch_bottom(const ocu_grid3d *u, const ocu_grid3d *b, const ocu_mg_level_params *p, const double *h, int nlev,
                  const ocu_bc_set *bc, int nu1, int nu2, int coarse_iters, int zero_top, hipStream_t stream) {
  BotArgs<T, TC> A;
  memset(&A, 0, sizeof(A));
  const size_t bytes = bottom_layout<T, TC>(u, nlev, MODE, &A);
  for (int l = 0; l < nlev; l++) {
    BotLevel<T> &L = A.L[l];
    L.u = dgrid<T>(&u[l]);
    L.b = dgrid<T>(&b[l]);
    L.dc = dev_c<T>(&p[l]);
    L.rc = res_c<T>(p[l].h, p[l].fx, p[l].fy, p[l].fz);
    L.gw = ghost_l1<T>(bc, h[3 * l], h[3 * l + 1], h[3 * l + 2]);
  }
  A.nlev = nlev;
  A.nu1 = nu1;
  A.nu2 = nu2;
  A.zero_top = zero_top;
  A.hc = host_c<TC>(&p[nlev - 1]);
  A.gwc = ghost_l1<TC>(bc, h[3 * (nlev - 1)], h[3 * (nlev - 1) + 1], h[3 * (nlev - 1) + 2]);
  A.per.a[0] = bc->xpos.type == OCU_BC_PERIODIC;
  A.per.a[1] = bc->ypos.type == OCU_BC_PERIODIC;
  A.per.a[2] = bc->zpos.type == OCU_BC_PERIODIC;
  A.coarse_iters = coarse_iters;
  // dynamic LDS beyond 64 KiB is requested explicitly
  OCU_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_mg_bottom<T, TC, MODE>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  k_mg_bottom<T, TC, MODE><<<1, BOT_THREADS, bytes, stream>>>(A);
  OCU_HIP_LAUNCHED("ocu_mg_bottom_vcycle");
}

}  // namespace

extern "C" {

size_t ocu_mg_bottom_lds_max(void) { return BOT_LDS_MAX; }

size_t ocu_mg_bottom_lds_bytes(const ocu_grid3d *u, int nlev, int coarse_mode) {
  if (!u || nlev < 2 || nlev > BOT_MAX) return 0;
  const bool dbl = u[0].dtype == OCU_F64;
  if (dbl) return bottom_layout<double, double>(u, nlev, 0, (BotArgs<double, double> *)nullptr);
  return coarse_mode == 1 ? bottom_layout<float, double>(u, nlev, 1, (BotArgs<float, double> *)nullptr)
                          : bottom_layout<float, float>(u, nlev, 0, (BotArgs<float, float> *)nullptr);
}

int ocu_mg_bottom_vcycle(const ocu_grid3d *u, const ocu_grid3d *b, const ocu_mg_level_params *p, const double *h,
                         int nlev, const ocu_bc_set *bc, int nu1, int nu2, int coarse_iters, int coarse_mode,
                         int zero_top, void *stream) {
  const char *what = "ocu_mg_bottom_vcycle";
  OCU_REQUIRE(u && b && p && h && bc, "%s: null argument", what);
  OCU_REQUIRE(nlev >= 2 && nlev <= BOT_MAX, "%s: %d levels (2..%d)", what, nlev, BOT_MAX);
  OCU_REQUIRE(nu1 >= 0 && nu2 >= 0 && coarse_iters >= 0, "%s: negative iteration count", what);
  OCU_REQUIRE(coarse_mode == 0 || (coarse_mode == 1 && u[0].dtype == OCU_F32), "%s: bad coarse mode %d", what,
              coarse_mode);
  for (int l = 0; l < nlev; l++) {
    if (int e = check_pair(&u[l], &b[l], what)) return e;
    OCU_REQUIRE(u[l].dtype == u[0].dtype, "%s: mixed dtypes", what);
    if (l > 0)
      OCU_REQUIRE(u[l - 1].nx >= 2 * u[l].nx && u[l - 1].ny >= 2 * u[l].ny && u[l - 1].nz >= 2 * u[l].nz &&
                      u[l].nx >= 1 && u[l].ny >= 1 && u[l].nz >= 1,
                  "%s: level %d is not a coarsening of level %d", what, l, l - 1);
  }
  if (!gfaces_paired(bc)) {
    set_error("%s: needs Neumann or Dirichlet faces or periodic face pairs", what);
    return OCU_ERR_UNSUPPORTED;
  }
  {
    const int per[3] = {bc->xpos.type == OCU_BC_PERIODIC, bc->ypos.type == OCU_BC_PERIODIC,
                        bc->zpos.type == OCU_BC_PERIODIC};
    for (int l = 0; l < nlev; l++) {
      const int n[3] = {u[l].nx, u[l].ny, u[l].nz};
      for (int a = 0; a < 3; a++)
        if (per[a] && n[a] % 2) {
          set_error("%s: periodic axis %d of level %d has an odd number of cells (%d)", what, a, l, n[a]);
          return OCU_ERR_UNSUPPORTED;
        }
    }
  }
  OCU_REQUIRE(!zero_top || gfaces_homogeneous(bc), "%s: zero start needs homogeneous faces", what);
  const size_t bytes = ocu_mg_bottom_lds_bytes(u, nlev, coarse_mode);
  OCU_REQUIRE(bytes > 0 && bytes <= BOT_LDS_MAX, "%s: %zu bytes of LDS exceed %zu", what, bytes, BOT_LDS_MAX);
  if (u[0].dtype == OCU_F64)
    return launch_bottom<double, double, 0>(u, b, p, h, nlev, bc, nu1, nu2, coarse_iters, zero_top, S(stream));
  if (coarse_mode == 1)
    return launch_bottom<float, double, 1>(u, b, p, h, nlev, bc, nu1, nu2, coarse_iters, zero_top, S(stream));
  return launch_bottom<float, float, 0>(u, b, p, h, nlev, bc, nu1, nu2, coarse_iters, zero_top, S(stream));
}


int ocu_mg_relax(const ocu_grid3d *u, const ocu_grid3d *b, int color, const ocu_mg_level_params *p,
                 int variant, void *stream) {
  if (int e = check_pair(u, b, "ocu_mg_relax")) return e;
  OCU_REQUIRE(p, "ocu_mg_relax: null params");
  OCU_REQUIRE(color == 0 || color == 1, "ocu_mg_relax: bad color %d", color);
  if ((int64_t)u->nx * u->ny * u->nz == 0) return OCU_OK;
  dim3 grid = stencil_grid(u->nx, u->ny, u->nz);
  if (variant == OCU_RELAX_DEVICE) {
    if (u->dtype == OCU_F64)
      k_relax_dev<double><<<grid, stencil_block(), 0, S(stream)>>>(dgrid<double>(u), dgrid<double>(b), dev_c64(p), color,
                                                                   GhostP<double>{}, 0);
    else
      k_relax_dev<float><<<grid, stencil_block(), 0, S(stream)>>>(dgrid<float>(u), dgrid<float>(b), dev_c32(p), color,
                                                                  GhostP<float>{}, 0);
  } else if (variant == OCU_RELAX_HOST) {
    if (u->dtype == OCU_F64)
      k_relax_host<double><<<grid, stencil_block(), 0, S(stream)>>>(dgrid<double>(u), dgrid<double>(b), host_c<double>(p), color);
    else
      k_relax_host<float><<<grid, stencil_block(), 0, S(stream)>>>(dgrid<float>(u), dgrid<float>(b), host_c<float>(p), color);
  } else {
    set_error("ocu_mg_relax: bad variant %d", variant);
    return OCU_ERR_INVALID;
  }
  OCU_HIP_LAUNCHED("ocu_mg_relax");
}

int ocu_mg_relax_bc(const ocu_grid3d *u, const ocu_grid3d *b, int color, const ocu_mg_level_params *p,
                    const ocu_bc_set *bc, double hx, double hy, double hz, void *stream) {
  const char *what = "ocu_mg_relax_bc";
  if (int e = check_pair(u, b, what)) return e;
  OCU_REQUIRE(p && bc, "%s: null argument", what);
  OCU_REQUIRE(color == 0 || color == 1, "%s: bad color %d", what, color);
  if (!gfaces_paired(bc)) {
    set_error("%s: ghost writes need Neumann or Dirichlet faces or periodic face pairs", what);
    return OCU_ERR_UNSUPPORTED;
  }
  const int n[3] = {u->nx, u->ny, u->nz};
  const int per[3] = {bc->xpos.type == OCU_BC_PERIODIC, bc->ypos.type == OCU_BC_PERIODIC,
                      bc->zpos.type == OCU_BC_PERIODIC};
  for (int a = 0; a < 3; a++)
    if (per[a] && n[a] % 2) {
      set_error("%s: a periodic axis needs an even number of cells (%d)", what, n[a]);
      return OCU_ERR_UNSUPPORTED;
    }
  if ((int64_t)u->nx * u->ny * u->nz == 0) return OCU_OK;
  dim3 grid = stencil_grid(u->nx, u->ny, u->nz);
  if (u->dtype == OCU_F64)
    k_relax_dev<double><<<grid, stencil_block(), 0, S(stream)>>>(dgrid<double>(u), dgrid<double>(b), dev_c64(p), color,
                                                                 ghost_p<double>(bc, hx, hy, hz), 1);
  else
    k_relax_dev<float><<<grid, stencil_block(), 0, S(stream)>>>(dgrid<float>(u), dgrid<float>(b), dev_c32(p), color,
                                                                ghost_p<float>(bc, hx, hy, hz), 1);
  OCU_HIP_LAUNCHED(what);
}

int ocu_mg_relax_fused(const ocu_grid3d *u_in, const ocu_grid3d *u_out, const ocu_grid3d *b,
                       const ocu_mg_level_params *p, const ocu_bc_set *bc, int first_color, void *stream) {
  return relax_fused(u_in, u_out, b, p, bc, 1.0, 1.0, 1.0, first_color, 0, stream);
}

int ocu_mg_relax_fused_bc(const ocu_grid3d *u_in, const ocu_grid3d *u_out, const ocu_grid3d *b,
                          const ocu_mg_level_params *p, const ocu_bc_set *bc, double hx, double hy, double hz,
                          int first_color, int flags, void *stream) {
  return relax_fused(u_in, u_out, b, p, bc, hx, hy, hz, first_color, flags, stream);
}

int ocu_mg_relax_fused_ok(const ocu_grid3d *u_in, const ocu_grid3d *u_out, const ocu_grid3d *b,
                          const ocu_bc_set *bc) {
  if (!u_in || !u_out || !b || !bc || check_grid(u_in, "ocu_mg_relax_fused_ok") ||
      check_grid(u_out, "ocu_mg_relax_fused_ok") || check_grid(b, "ocu_mg_relax_fused_ok"))
    return 0;
  if (!any_periodic(bc)) return gfaces_local(bc) ? 1 : 0;
  return fused_periodic_ok(u_in, u_out, b, bc);
}

int ocu_mg_relax_fused_prolong_ok(const ocu_grid3d *u, const ocu_grid3d *u_coarse, const ocu_bc_set *bc) {
  if (!u || !u_coarse || check_grid(u, "ocu_mg_relax_fused_prolong_ok") ||
      check_grid(u_coarse, "ocu_mg_relax_fused_prolong_ok"))
    return 0;
  return fused_prolong_width(u, u, u, u_coarse, bc) ? 1 : 0;
}

int ocu_mg_relax_fused_prolong_bc(const ocu_grid3d *u_in, const ocu_grid3d *u_out, const ocu_grid3d *b,
                                  const ocu_grid3d *u_coarse, int quirk, const ocu_mg_level_params *p,
                                  const ocu_bc_set *bc, double hx, double hy, double hz, int first_color, int flags,
                                  void *stream) {
  const char *what = "ocu_mg_relax_fused_prolong";
  if (int e = check_pair(u_in, b, what)) return e;
  if (int e = check_grid(u_out, "ocu_mg_relax_fused_prolong(out)")) return e;
  if (int e = check_grid(u_coarse, "ocu_mg_relax_fused_prolong(coarse)")) return e;
  OCU_REQUIRE(layout_match(u_in, u_out) && u_in->dtype == u_out->dtype && u_in->nx == u_out->nx &&
                  u_in->ny == u_out->ny && u_in->nz == u_out->nz,
              "%s: u_out must be congruent with u_in", what);
  OCU_REQUIRE(u_out->gx >= 1 && u_out->gy >= 1 && u_out->gz >= 1, "%s: u_out needs one ghost layer", what);
  OCU_REQUIRE((flags & ~OCU_FUSED_WRITE_GHOSTS) == 0, "%s: bad flags %d", what, flags);
  OCU_REQUIRE(u_in->base != u_out->base && u_coarse->base != u_out->base, "%s: u_out must not alias an input",
              what);
  OCU_REQUIRE(p && bc, "%s: null params", what);
  OCU_REQUIRE(first_color == 0 || first_color == 1, "%s: bad first_color %d", what, first_color);
  const int V = fused_prolong_width(u_in, u_out, b, u_coarse, bc);
  if (!V) {
    set_error("%s: needs fine = 2 x coarse, whole rows of 4 cells per lane, aligned coarse rows and "
              "Neumann or Dirichlet faces", what);
    return OCU_ERR_UNSUPPORTED;
  }
  const int wghost = (flags & OCU_FUSED_WRITE_GHOSTS) ? 1 : 0;
  if (u_in->dtype == OCU_F64) {
    const DevC64 c = dev_c64(p);
    const GhostL1<double> gw = ghost_l1<double>(bc, hx, hy, hz);
    return launch_fused_rows<double, 4, false, true>(u_in, u_out, b, c, first_color, gw, wghost, S(stream),
                                                     u_coarse, quirk);
  }
  const DevC32 c = dev_c32(p);
  const GhostL1<float> gw = ghost_l1<float>(bc, hx, hy, hz);
  return launch_fused_rows<float, 4, false, true>(u_in, u_out, b, c, first_color, gw, wghost, S(stream), u_coarse,
                                                  quirk);
}

int ocu_mg_residual(const ocu_grid3d *u, const ocu_grid3d *b, const ocu_grid3d *r, double h, double fx,
                    double fy, double fz, void *stream) {
  if (int e = check_pair(u, b, "ocu_mg_residual")) return e;
  if (int e = check_grid(r, "ocu_mg_residual(r)")) return e;
  OCU_REQUIRE(r->dtype == u->dtype && r->nx == u->nx && r->ny == u->ny && r->nz == u->nz,
              "ocu_mg_residual: r mismatch");
  if ((int64_t)u->nx * u->ny * u->nz == 0) return OCU_OK;
  dim3 grid = stencil_grid(u->nx, u->ny, u->nz);
  if (u->dtype == OCU_F64)
    k_residual<double><<<grid, stencil_block(), 0, S(stream)>>>(dgrid<double>(u), dgrid<double>(b), dgrid<double>(r),
                                                                res_c<double>(h, fx, fy, fz));
  else
    k_residual<float><<<grid, stencil_block(), 0, S(stream)>>>(dgrid<float>(u), dgrid<float>(b), dgrid<float>(r),
                                                               res_c<float>(h, fx, fy, fz));
  OCU_HIP_LAUNCHED("ocu_mg_residual");
}

int ocu_mg_restrict(const ocu_grid3d *r, const ocu_grid3d *b_coarse, void *stream) {
  if (int e = check_grid(r, "ocu_mg_restrict(r)")) return e;
  if (int e = check_grid(b_coarse, "ocu_mg_restrict(b)")) return e;
  OCU_REQUIRE(r->dtype == b_coarse->dtype && r->dtype != OCU_I32, "ocu_mg_restrict: dtype mismatch");
  OCU_REQUIRE(r->nx >= 2 * b_coarse->nx && r->ny >= 2 * b_coarse->ny && r->nz >= 2 * b_coarse->nz,
              "ocu_mg_restrict: fine grid must be twice the coarse grid");
  if ((int64_t)b_coarse->nx * b_coarse->ny * b_coarse->nz == 0) return OCU_OK;
  dim3 grid = stencil_grid(b_coarse->nx, b_coarse->ny, b_coarse->nz);
  if (r->dtype == OCU_F64)
    k_restrict<double><<<grid, stencil_block(), 0, S(stream)>>>(dgrid<double>(r), dgrid<double>(b_coarse));
  else
    k_restrict<float><<<grid, stencil_block(), 0, S(stream)>>>(dgrid<float>(r), dgrid<float>(b_coarse));
  OCU_HIP_LAUNCHED("ocu_mg_restrict");
}

int ocu_mg_residual_restrict(const ocu_grid3d *u, const ocu_grid3d *b, const ocu_grid3d *r,
                             const ocu_grid3d *b_coarse, double h, double fx, double fy, double fz,
                             double *norm_ws, void *stream) {
  if (int e = check_pair(u, b, "ocu_mg_residual_restrict")) return e;
  if (int e = check_grid(b_coarse, "ocu_mg_residual_restrict(b_coarse)")) return e;
  OCU_REQUIRE(b_coarse->dtype == u->dtype, "ocu_mg_residual_restrict: dtype mismatch");
  OCU_REQUIRE(u->nx == 2 * b_coarse->nx && u->ny == 2 * b_coarse->ny && u->nz == 2 * b_coarse->nz,
              "ocu_mg_residual_restrict: fine grid must be exactly twice the coarse grid");
  if (r) {
    if (int e = check_grid(r, "ocu_mg_residual_restrict(r)")) return e;
    OCU_REQUIRE(r->dtype == u->dtype && r->nx == u->nx && r->ny == u->ny && r->nz == u->nz,
                "ocu_mg_residual_restrict: r mismatch");
  }
  const int cnx = b_coarse->nx, cny = b_coarse->ny, cnz = b_coarse->nz;
  if ((int64_t)cnx * cny * cnz == 0) return OCU_OK;
  if (!r) {
    switch (res_row_width(u, b, b_coarse)) {
      case 2: return launch_res_restrict_rows<2>(u, b, b_coarse, h, fx, fy, fz, norm_ws, S(stream));
      case 4: return launch_res_restrict_rows<4>(u, b, b_coarse, h, fx, fy, fz, norm_ws, S(stream));
      default: break;
    }
  }
  int tiles_k = (cnz + TK - 1) / TK, tiles_j = (cny + TJ - 1) / TJ;
  int ntiles = tiles_k * tiles_j * cnx;
  dim3 block = stencil_block();
  bool wr_r = r != nullptr;
#define RR_LAUNCH(T, NORM, NB)                                                                          \
  k_res_restrict<T, NORM><<<NB, block, 0, S(stream)>>>(dgrid<T>(u), dgrid<T>(b), wr_r ? dgrid<T>(r) : dgrid<T>(u), \
                                                       wr_r, dgrid<T>(b_coarse), res_c<T>(h, fx, fy, fz), tiles_k,   \
                                                       tiles_j, ntiles, norm_ws)
  if (norm_ws) {
    // always NORM_PARTS blocks: idle blocks write zero partials, so no
    // memset is needed (keeps the call safe inside a captured graph)
    if (u->dtype == OCU_F64) RR_LAUNCH(double, true, NORM_PARTS);
    else RR_LAUNCH(float, true, NORM_PARTS);
  } else {
    if (u->dtype == OCU_F64) RR_LAUNCH(double, false, ntiles);
    else RR_LAUNCH(float, false, ntiles);
  }
#undef RR_LAUNCH
  OCU_HIP_LAUNCHED("ocu_mg_residual_restrict");
}

int ocu_mg_residual_norm(const ocu_grid3d *u, const ocu_grid3d *b, const ocu_grid3d *r, double h, double fx,
                         double fy, double fz, double *norm_ws, void *stream) {
  if (!norm_ws) return ocu_mg_residual(u, b, r, h, fx, fy, fz, stream);
  if (int e = check_pair(u, b, "ocu_mg_residual_norm")) return e;
  if (int e = check_grid(r, "ocu_mg_residual_norm(r)")) return e;
  OCU_REQUIRE(r->dtype == u->dtype && r->nx == u->nx && r->ny == u->ny && r->nz == u->nz,
              "ocu_mg_residual_norm: r mismatch");
  int tiles_k = (u->nz + TK - 1) / TK, tiles_j = (u->ny + TJ - 1) / TJ;
  int ntiles = tiles_k * tiles_j * u->nx;
  const int nb = NORM_PARTS;
  if (u->dtype == OCU_F64)
    k_res_norm<double><<<nb, stencil_block(), 0, S(stream)>>>(dgrid<double>(u), dgrid<double>(b), dgrid<double>(r),
                                                              res_c<double>(h, fx, fy, fz), tiles_k, tiles_j, ntiles, norm_ws);
  else
    k_res_norm<float><<<nb, stencil_block(), 0, S(stream)>>>(dgrid<float>(u), dgrid<float>(b), dgrid<float>(r),
                                                             res_c<float>(h, fx, fy, fz), tiles_k, tiles_j, ntiles, norm_ws);
  OCU_HIP_LAUNCHED("ocu_mg_residual_norm");
}

int ocu_norm_finalize(const double *norm_ws, double *out, void *stream) {
  OCU_REQUIRE(norm_ws && out, "ocu_norm_finalize: null pointer");
  k_norm_finalize<<<1, 512, 0, S(stream)>>>(norm_ws, NORM_PARTS, out, 0ull);
  OCU_HIP_LAUNCHED("ocu_norm_finalize");
}

int ocu_norm_finalize_tagged(const double *norm_ws, double *out, unsigned long long tag, void *stream) {
  OCU_REQUIRE(norm_ws && out, "ocu_norm_finalize_tagged: null pointer");
  OCU_REQUIRE(tag, "ocu_norm_finalize_tagged: the tag must not be 0");
  k_norm_finalize<<<1, 512, 0, S(stream)>>>(norm_ws, NORM_PARTS, out, tag);
  OCU_HIP_LAUNCHED("ocu_norm_finalize_tagged");
}

int ocu_mg_prolong(const ocu_grid3d *u_coarse, const ocu_grid3d *u_fine, int quirk, void *stream) {
  return prolong(u_coarse, u_fine, quirk, nullptr, 1.0, 1.0, 1.0, stream);
}

int ocu_mg_prolong_bc_ok(const ocu_grid3d *u_coarse, const ocu_grid3d *u_fine, const ocu_bc_set *bc) {
  if (!u_coarse || !u_fine || !bc || check_grid(u_coarse, "ocu_mg_prolong_bc_ok") ||
      check_grid(u_fine, "ocu_mg_prolong_bc_ok"))
    return 0;
  if (u_fine->nx != 2 * u_coarse->nx || u_fine->ny != 2 * u_coarse->ny || u_fine->nz != 2 * u_coarse->nz) return 0;
  if (u_fine->gx < 1 || u_fine->gy < 1 || u_fine->gz < 1 || u_coarse->gx < 1 || u_coarse->gy < 1 ||
      u_coarse->gz < 1)
    return 0;
  return prolong_bc_ok(u_coarse, u_fine, bc) ? 1 : 0;
}

int ocu_mg_prolong_bc(const ocu_grid3d *u_coarse, const ocu_grid3d *u_fine, int quirk, const ocu_bc_set *bc,
                      double hx, double hy, double hz, void *stream) {
  OCU_REQUIRE(bc, "ocu_mg_prolong_bc: null bc");
  return prolong(u_coarse, u_fine, quirk, bc, hx, hy, hz, stream);
}

// ocu_mg_coarse_relax for boxes beyond one workgroup's LDS
static int coarse_relax_global(const ocu_grid3d *u, const ocu_grid3d *b, CoarseGeom g, size_t box, int iterations,
                               int order, const ocu_mg_level_params *p, const FacesS &F, double hx, double hy,
                               double hz, int mode, void *stream) {
  OCU_REQUIRE(box < (1u << 30), "ocu_mg_coarse_relax: grid %dx%dx%d too large", g.nx, g.ny, g.nz);
  if (mode == 1) {
    double *w = nullptr;
    OCU_HIP_TRY(hipMallocAsync((void **)&w, 2 * box * sizeof(double), S(stream)));
    k_coarse_global<float, double, 1><<<1, 1024, 0, S(stream)>>>(dgrid<float>(u), dgrid<float>(b), w, w + box, g,
                                                                 host_c<double>(p), F, hx, hy, hz, iterations, order);
    const hipError_t le = hipGetLastError();
    OCU_HIP_TRY(hipFreeAsync(w, S(stream)));
    OCU_HIP_TRY(le);
    return OCU_OK;
  }
  OCU_REQUIRE(layout_match(u, b), "ocu_mg_coarse_relax: u and b layouts differ (needed beyond LDS size)");
  const int64_t xs = (int64_t)u->pny * u->pnz;
  OCU_REQUIRE(xs < (1ll << 30), "ocu_mg_coarse_relax: grid %dx%dx%d too large", g.nx, g.ny, g.nz);
  g.lxs = (int)xs;
  g.lys = u->pnz;
  g.lshift = 0;
  if (u->dtype == OCU_F64) {
    DGrid<double> du = dgrid<double>(u), db = dgrid<double>(b);
    k_coarse_global<double, double, 0><<<1, 1024, 0, S(stream)>>>(du, db, du.p, db.p, g, host_c<double>(p), F, hx,
                                                                  hy, hz, iterations, order);
  } else {
    DGrid<float> du = dgrid<float>(u), db = dgrid<float>(b);
    k_coarse_global<float, float, 0><<<1, 1024, 0, S(stream)>>>(du, db, du.p, db.p, g, host_c<float>(p), F,
                                                                (float)hx, (float)hy, (float)hz, iterations, order);
  }
  OCU_HIP_LAUNCHED("ocu_mg_coarse_relax");
}

int ocu_mg_coarse_relax(const ocu_grid3d *u, const ocu_grid3d *b, int iterations, int order,
                        const ocu_mg_level_params *p, const ocu_bc_set *bc, double hx, double hy, double hz,
                        int mode, void *stream) {
  if (int e = check_pair(u, b, "ocu_mg_coarse_relax")) return e;
  OCU_REQUIRE(p && bc, "ocu_mg_coarse_relax: null params");
  OCU_REQUIRE(order >= 0 && order <= 2, "ocu_mg_coarse_relax: bad order %d", order);
  OCU_REQUIRE(mode == 0 || (mode == 1 && u->dtype == OCU_F32), "ocu_mg_coarse_relax: bad mode %d", mode);
  if (order == 2) {
    if (iterations % 2 == 1) iterations++;
    if (iterations < 4) iterations = 4;
  }
  CoarseGeom g;
  g.nx = u->nx; g.ny = u->ny; g.nz = u->nz;
  g.lys = g.nz + 2;
  g.lxs = (g.ny + 2) * g.lys;
  g.lshift = g.lxs + g.lys + 1;
  size_t box = (size_t)(g.nx + 2) * (g.ny + 2) * (g.nz + 2);
  size_t tsize = (mode == 1 || u->dtype == OCU_F64) ? 8 : 4;
  size_t shmem = 2 * box * tsize;
  FacesS F = faces_s(bc);
  if (shmem > BOT_LDS_MAX) return coarse_relax_global(u, b, g, box, iterations, order, p, F, hx, hy, hz, mode, stream);
  // dynamic LDS beyond 64 KiB is requested explicitly
  const void *kc = mode == 1 ? reinterpret_cast<const void *>(&k_coarse<float, double, 1>)
                   : u->dtype == OCU_F64 ? reinterpret_cast<const void *>(&k_coarse<double, double, 0>)
                                         : reinterpret_cast<const void *>(&k_coarse<float, float, 0>);
  OCU_HIP_TRY(hipFuncSetAttribute(kc, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
  if (mode == 1) {
    k_coarse<float, double, 1><<<1, 256, shmem, S(stream)>>>(dgrid<float>(u), dgrid<float>(b), g, host_c<double>(p), F,
                                                             hx, hy, hz, iterations, order);
  } else if (u->dtype == OCU_F64) {
    k_coarse<double, double, 0><<<1, 256, shmem, S(stream)>>>(dgrid<double>(u), dgrid<double>(b), g, host_c<double>(p),
                                                              F, hx, hy, hz, iterations, order);
  } else {
    k_coarse<float, float, 0><<<1, 256, shmem, S(stream)>>>(dgrid<float>(u), dgrid<float>(b), g, host_c<float>(p), F,
                                                            (float)hx, (float)hy, (float)hz, iterations, order);
  }
  OCU_HIP_LAUNCHED("ocu_mg_coarse_relax");
}

}  // extern "C"
