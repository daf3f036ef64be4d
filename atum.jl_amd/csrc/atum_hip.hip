// libatumhip.so: C-ABI implementation (contexts, device memory, DGSEM operator,
// fused LSRK driver with hipGraph replay, banded vertical solver, ARK23).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "handles.hpp"

using namespace atm;

struct atm_batch {
  atm_ctx* ctx = nullptr;
  // a device table and its pinned host copy (the source of its uploads, kept for
  // the batch's life), with an event after the last upload from it; a slot keeps
  // its memory when the members change and grows only when a group outgrows it
  struct Slot {
    BatchEntry* dtab = nullptr;
    BatchEntry* htab = nullptr;
    hipEvent_t uploaded = nullptr;
    size_t cap = 0;      // entries
    size_t used = 0;     // entries of the table it holds
    bool valid = false;  // htab/dtab hold that table
  };
  struct Group {
    BatchLaunchFn bfn;
    int Np, epw, threads;
    size_t lds;
    std::vector<int> members;
    int slot = -1;
    int64_t blocks = 0;
    // members of different Nq in one launch: each keeps the shape it had when the
    // batch was created (threads and lds are the maxima over the members)
    bool multi = false;
    std::vector<int> mepw;
    std::vector<int> mlds;
  };
  std::vector<atm_dg*> dgs;
  std::vector<Group> groups;
  std::vector<Slot> slots;
  hipGraphExec_t exec[2] = {nullptr, nullptr};  // two-step stage sequence per scheme (reads the tables)
};

struct atm_banded {
  atm_dg* lin = nullptr;
  int n = 0, bw = 0, nb = 0, KV = 0, ncols = 0, Nh = 0;
  int64_t nlines = 0;
  int64_t nu = 0;               // distinct line matrices
  double* fac = nullptr;        // LU factors of the distinct matrices: [u * n*nb + r*nb + bw + c - r]
  int32_t* line_mat = nullptr;  // matrix of each line
  std::vector<int32_t> line_mat_host;
  int32_t* perm = nullptr;      // lines ordered by matrix, each matrix's run padded to 64 with -1
  int32_t* wave_mat = nullptr;  // matrix of each run of 64
  int64_t nwaves = 0;
  // block form of the factors (element blocks of BS = bw+1 rows): per matrix u and block i,
  // [L_{i,i-1} | L_ii^-1 | U_{i,i+1} | U_ii^-1], each BS x BS row-major (null: not built)
  double* blocks = nullptr;
  double* frags = nullptr;        // the block form as MFMA A-operand fragments (band_frags_kernel)
  int32_t* perm16 = nullptr;      // as perm, runs of 16 lines; padding repeats the run's first line
  int32_t* wave_mat16 = nullptr;
  int64_t nwaves16 = 0;
  int32_t* elem_col = nullptr;  // column of each element (-1: none)
  int32_t* elem_ez = nullptr;   // position of each element in its column
  int32_t* colelem = nullptr;   // ncols*KV
  size_t dev_bytes = 0;         // device memory held (factors, block form, index tables)
  int kernel = ATM_BAND_AUTO;   // solve kernel (ATM_BAND_*)
  int mfma_w = 0;               // waves per workgroup of the matrix-core solve (0: the largest that fits)
};

namespace atm {
thread_local hipStream_t tl_stream = nullptr;
atm_dg* banded_lin(const atm_banded* b) { return b ? b->lin : nullptr; }
}

namespace {

constexpr int EW_THREADS = 256;

inline int ew_blocks(int64_t n) {
  int64_t b = (n + EW_THREADS - 1) / EW_THREADS;
  return (int)(b > 65535 * 16 ? 65535 * 16 : (b < 1 ? 1 : b));
}

__global__ void axpby_kernel(double* __restrict__ y, const double* __restrict__ x, double a, double b,
                             int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = a * y[i] + b * x[i];
}

__global__ void fill_kernel(double* __restrict__ y, double v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = v;
}

__global__ void copy_kernel(double* __restrict__ y, const double* __restrict__ x, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = x[i];
}


struct ArkCoef {
  double ea[3], ia[3];
};

// The right-hand side of an implicit ARK stage, q + dt sum_k<nk (ea_k ex_k + ia_k im_k), formed
// where it is read (the banded solve's row loads) instead of in a pass of its own; the same
// operations in the same order as ark_combine_kernel, so the values are the same bits.
struct ArkRhs {
  const double* q;
  const double* e[3];
  const double* i[3];
  ArkCoef c;
  double dt;
  int unsplit;
};

// Qhat = q; Qhat += dt*(a_k*exK_k + ã_k*imK_k) for k < nk (src/odesolvers.jl:346-349).
// unsplit: exK_k holds the full tendency and the explicit part is exK_k - imK_k
// (src/odesolvers.jl:339-341, 374-376), formed here with the same rounding.
__global__ void ark_combine_kernel(double* out, const double* q, double dt,
                                   int nk, ArkCoef c, const double* __restrict__ e0,
                                   const double* __restrict__ e1, const double* __restrict__ e2,
                                   const double* __restrict__ i0, const double* __restrict__ i1,
                                   const double* __restrict__ i2, int64_t n, int unsplit) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double v = q[i];
    if (nk > 0) {
      const double im = i0[i], ex = unsplit ? e0[i] - im : e0[i];
      v = v + dt * (c.ea[0] * ex + c.ia[0] * im);
    }
    if (nk > 1) {
      const double im = i1[i], ex = unsplit ? e1[i] - im : e1[i];
      v = v + dt * (c.ea[1] * ex + c.ia[1] * im);
    }
    if (nk > 2) {
      const double im = i2[i], ex = unsplit ? e2[i] - im : e2[i];
      v = v + dt * (c.ea[2] * ex + c.ia[2] * im);
    }
    out[i] = v;
  }
}

// ---- banded -----------------------------------------------------------------
struct BandGeo {
  int Ns, Nq, Nh, Np, KV, n, bw, nb;
  int64_t nlines, Ntot;
  const int32_t* elem_col;
  const int32_t* elem_ez;
  const int32_t* colelem;
};

__device__ __forceinline__ int64_t band_node(const BandGeo& g, int64_t line, int r, int* s) {
  const int64_t c = line / g.Nh;
  const int h = (int)(line - c * g.Nh);
  *s = r % g.Ns;
  const int rem = r / g.Ns;
  const int k = rem % g.Nq, ez = rem / g.Nq;
  const int64_t e = g.colelem[c * g.KV + ez];
  return (int64_t)h + (int64_t)g.Nh * k + (int64_t)g.Np * e;
}

// probe vector: 1 on every unknown whose in-line index m satisfies m % nb == color
__global__ void band_probe_kernel(BandGeo g, double* __restrict__ q, int color) {
  const int64_t tot = (int64_t)g.Ns * g.Ntot;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
    const int s = (int)(t / g.Ntot);
    const int64_t node = t - (int64_t)s * g.Ntot;
    const int64_t e = node / g.Np;
    const int ijk = (int)(node - e * g.Np);
    double v = 0.0;
    if (g.elem_col[e] >= 0) {
      const int k = ijk / g.Nh;
      const int m = s + g.Ns * (k + g.Nq * g.elem_ez[e]);
      v = (m % g.nb == color) ? 1.0 : 0.0;
    }
    q[t] = v;
  }
}

__global__ void band_scatter_kernel(BandGeo g, const double* __restrict__ dq, double* __restrict__ band,
                                    int color) {
  const int64_t tot = g.nlines * g.n;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(t / g.nlines);
    const int64_t line = t - (int64_t)r * g.nlines;
    const int lo = r - g.bw > 0 ? r - g.bw : 0;
    const int m = lo + ((color - lo) % g.nb + g.nb) % g.nb;
    const int hi = r + g.bw < g.n - 1 ? r + g.bw : g.n - 1;
    if (m > hi) continue;
    int s;
    const int64_t node = band_node(g, line, r, &s);
    band[((int64_t)r * g.nb + (g.bw + m - r)) * g.nlines + line] = dq[(int64_t)s * g.Ntot + node];
  }
}

// A = I + scale*L over the assembled band (layout [(r*nb + j) * nlines + line]).
__global__ void band_scale_kernel(BandGeo g, double* __restrict__ band, double scale) {
  const int64_t tot = (int64_t)g.n * g.nb * g.nlines;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
    const int j = (int)((t / g.nlines) % g.nb);
    double v = band[t] * scale;
    if (j == g.bw) v += 1;
    band[t] = v;
  }
}

// Line hashes for finding lines with bit-identical bands: 64-bit FNV-1a over
// the bit patterns of each chunk of HASH_CHUNK band entries (one thread per
// line and chunk), then FNV-1a over a line's chunk hashes in chunk order.
constexpr int HASH_CHUNK = 64;

__global__ void band_hash_chunk_kernel(BandGeo g, const double* __restrict__ band, uint64_t* __restrict__ part) {
  const int64_t line = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (line >= g.nlines) return;
  const int64_t m = (int64_t)g.n * g.nb;
  const int64_t k0 = (int64_t)blockIdx.y * HASH_CHUNK;
  const int64_t k1 = k0 + HASH_CHUNK < m ? k0 + HASH_CHUNK : m;
  uint64_t h = 1469598103934665603ull;
  for (int64_t k = k0; k < k1; ++k) {
    h ^= (uint64_t)__double_as_longlong(band[k * g.nlines + line]);
    h *= 1099511628211ull;
  }
  part[(int64_t)blockIdx.y * g.nlines + line] = h;
}

__global__ void band_hash_combine_kernel(BandGeo g, const uint64_t* __restrict__ part, int nchunk,
                                         uint64_t* __restrict__ hash) {
  const int64_t line = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (line >= g.nlines) return;
  uint64_t h = 1469598103934665603ull;
  for (int c = 0; c < nchunk; ++c) {
    h ^= part[(int64_t)c * g.nlines + line];
    h *= 1099511628211ull;
  }
  hash[line] = h;
}

// same[line] is set to 1 beforehand; a chunk that differs bit-wise from the
// same chunk of the representative line rep[line] clears it.
__global__ void band_verify_chunk_kernel(BandGeo g, const double* __restrict__ band, const int64_t* __restrict__ rep,
                                         int32_t* __restrict__ same) {
  const int64_t line = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (line >= g.nlines) return;
  const int64_t o = rep[line];
  const int64_t m = (int64_t)g.n * g.nb;
  const int64_t k0 = (int64_t)blockIdx.y * HASH_CHUNK;
  const int64_t k1 = k0 + HASH_CHUNK < m ? k0 + HASH_CHUNK : m;
  bool eq = true;
  for (int64_t k = k0; k < k1; ++k)
    eq = eq && __double_as_longlong(band[k * g.nlines + line]) == __double_as_longlong(band[k * g.nlines + o]);
  if (!eq) same[line] = 0;
}

__global__ void fill_i32_kernel(int32_t* __restrict__ y, int32_t v, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = v;
}

// fac[u * n*nb + k] = band[k * nlines + src[u]]
__global__ void band_compact_kernel(BandGeo g, const double* __restrict__ band, const int64_t* __restrict__ src,
                                    int64_t nu, double* __restrict__ fac) {
  const int64_t M = (int64_t)g.n * g.nb;
  const int64_t tot = M * nu;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t u = t / M, k = t - u * M;
    fac[t] = band[k * g.nlines + src[u]];
  }
}

// In-place LU without pivoting (unit lower factor), one workgroup per distinct
// matrix: at step k the threads form the multipliers of rows k+1..k+bw, then
// update the trailing (bw x bw) block.  Every entry sees the same operations
// in the same order as in the sequential elimination.  With LDS_COPY the
// matrix is staged in LDS ([r*nb + j]) for the elimination.
constexpr int LU_THREADS = 256;

template <bool LDS_COPY>
__global__ void __launch_bounds__(LU_THREADS) band_lu_wg_kernel(BandGeo g, double* __restrict__ fac, int64_t nu) {
  extern __shared__ double sA[];
  const int64_t u = blockIdx.x;
  const int n = g.n, bw = g.bw, nb = g.nb;
  const int t = threadIdx.x;
  const int64_t m = (int64_t)n * nb;
  double* Fu = fac + u * m;
  if (LDS_COPY)
    for (int64_t k = t; k < m; k += LU_THREADS) sA[k] = Fu[k];
  __syncthreads();
  auto A = [&](int r, int c) -> double& {
    const int64_t k = (int64_t)r * nb + (bw + c - r);
    return LDS_COPY ? sA[k] : Fu[k];
  };
  for (int k = 0; k < n; ++k) {
    const int iend = k + bw < n - 1 ? k + bw : n - 1;
    const int w = iend - k;
    if (t < w) {
      const int i = k + 1 + t;
      A(i, k) = A(i, k) / A(k, k);
    }
    __syncthreads();
    for (int q = t; q < w * w; q += LU_THREADS) {
      const int i = k + 1 + q / w, j = k + 1 + q % w;
      A(i, j) -= A(i, k) * A(k, j);
    }
    __syncthreads();
  }
  if (LDS_COPY)
    for (int64_t k = t; k < m; k += LU_THREADS) Fu[k] = sA[k];
}

// Forward and back substitution, one thread (lane) per vertical line; the last
// bw unknowns of a line live in an LDS window; unknowns are visited with
// incremental addressing (element, vertical node, state).  x may alias b.
constexpr int BAND_WIN = 64;  // > bw for every supported (Ns, Nq)

__global__ void __launch_bounds__(64)
band_solve_kernel(BandGeo g, const double* __restrict__ fac, const int32_t* __restrict__ line_mat, int64_t nu,
                  double* x, const double* b) {
  __shared__ double win[BAND_WIN][64];
  const int lane = threadIdx.x;
  const int64_t line = blockIdx.x * (int64_t)blockDim.x + lane;
  if (line >= g.nlines) return;
  const int n = g.n, bw = g.bw, nb = g.nb;
  const int64_t mat = line_mat[line];
  const int64_t c = line / g.Nh;
  const int h = (int)(line - c * g.Nh);
  const int32_t* col = g.colelem + c * g.KV;
  int r = 0;
  for (int ez = 0; ez < g.KV; ++ez) {
    const int64_t base_e = h + (int64_t)g.Np * col[ez];
    for (int k = 0; k < g.Nq; ++k) {
      const int64_t base = base_e + (int64_t)g.Nh * k;
      for (int s = 0; s < g.Ns; ++s, ++r) {
        const int64_t idx = (int64_t)s * g.Ntot + base;
        double v = b[idx];
        const int lo = r - bw > 0 ? r - bw : 0;
        const double* arow = fac + mat * ((int64_t)n * nb) + (int64_t)r * nb + bw - r;
        for (int cc = lo; cc < r; ++cc) v -= arow[cc] * win[cc & (BAND_WIN - 1)][lane];
        win[r & (BAND_WIN - 1)][lane] = v;
        x[idx] = v;
      }
    }
  }
  r = n - 1;
  for (int ez = g.KV - 1; ez >= 0; --ez) {
    const int64_t base_e = h + (int64_t)g.Np * col[ez];
    for (int k = g.Nq - 1; k >= 0; --k) {
      const int64_t base = base_e + (int64_t)g.Nh * k;
      for (int s = g.Ns - 1; s >= 0; --s, --r) {
        const int64_t idx = (int64_t)s * g.Ntot + base;
        double v = x[idx];
        const int hi = r + bw < n - 1 ? r + bw : n - 1;
        const double* arow = fac + mat * ((int64_t)n * nb) + (int64_t)r * nb + bw - r;
        for (int cc = r + 1; cc <= hi; ++cc) v -= arow[cc] * win[cc & (BAND_WIN - 1)][lane];
        v = v / arow[r];
        win[r & (BAND_WIN - 1)][lane] = v;
        x[idx] = v;
      }
    }
  }
}

// Forward and back substitution with the line vectors staged in LDS: a
// workgroup (one wave) takes 64/PARTS lines, PARTS lanes per line; each lane
// sums every PARTS-th product of a row's band window and the lanes of a line
// combine their partial sums by cross-lane exchange.  Loads and stores of the
// line vectors are coalesced over the lines' horizontal node index.
// LDS: y[r * (LINES + 1) + j].  x may alias b (all of a line is read first).
template <int PARTS>
__global__ void __launch_bounds__(64)
band_solve_lds_kernel(BandGeo g, const double* __restrict__ fac, const int32_t* __restrict__ line_mat, int64_t nu,
                      double* x, const double* b) {
  constexpr int LINES = 64 / PARTS, LD = LINES + 1;
  extern __shared__ double y[];
  const int t = threadIdx.x, j = t / PARTS, part = t % PARTS;
  const int64_t line = blockIdx.x * (int64_t)LINES + j;
  const bool active = line < g.nlines;
  const int n = g.n, bw = g.bw, nb = g.nb;
  const int64_t c = active ? line / g.Nh : 0;
  const int h = active ? (int)(line - c * g.Nh) : 0;
  const int32_t* col = g.colelem + c * g.KV;
  const int64_t mat = active ? line_mat[line] : 0;
  // unknown r = s + Ns*(k + Nq*ez) lives at s*Ntot + h + Nh*k + Np*col[ez]
  auto gidx = [&](int r) -> int64_t {
    const int s = r % g.Ns, rem = r / g.Ns;
    const int k = rem % g.Nq, ez = rem / g.Nq;
    return (int64_t)s * g.Ntot + h + (int64_t)g.Nh * k + (int64_t)g.Np * col[ez];
  };
  for (int r = part; r < n; r += PARTS) y[r * LD + j] = active ? b[gidx(r)] : 0.0;
  __syncthreads();
  for (int r = 0; r < n; ++r) {
    const int lo = r - bw > 0 ? r - bw : 0;
    const double* arow = fac + mat * ((int64_t)n * nb) + (int64_t)r * nb + bw - r;
    double acc = 0.0;
    for (int cc = lo + part; cc < r; cc += PARTS) acc += arow[cc] * y[cc * LD + j];
#pragma unroll
    for (int o = 1; o < PARTS; o <<= 1) acc += __shfl_xor(acc, o, 64);
    const double v = y[r * LD + j] - acc;
    __syncthreads();
    if (part == 0) y[r * LD + j] = v;
    __syncthreads();
  }
  for (int r = n - 1; r >= 0; --r) {
    const int hi = r + bw < n - 1 ? r + bw : n - 1;
    const double* arow = fac + mat * ((int64_t)n * nb) + (int64_t)r * nb + bw - r;
    double acc = 0.0;
    for (int cc = r + 1 + part; cc <= hi; cc += PARTS) acc += arow[cc] * y[cc * LD + j];
#pragma unroll
    for (int o = 1; o < PARTS; o <<= 1) acc += __shfl_xor(acc, o, 64);
    const double v = (y[r * LD + j] - acc) / arow[r];
    __syncthreads();
    if (part == 0) y[r * LD + j] = v;
    __syncthreads();
  }
  if (!active) return;
  for (int r = part; r < n; r += PARTS) x[gidx(r)] = y[r * LD + j];
}

// Forward and back substitution with one lane per line and the unknowns of
// one element of the line (BS = NS*NQ rows, the band half-width + 1) held in
// registers, fully unrolled: row r depends on the previous block's and the
// current block's values only, so the products of a row are independent
// except for the last few, and the rows of a block overlap (each block of rows
// is one branch-free sequence; the first and last blocks are separate
// instantiations).  Lines are taken in the order perm (grouped by matrix, each
// group padded to 64 with -1), so every wave solves with one matrix,
// wave_mat[wave].  The wave loads the next block's factor rows (BS x nb)
// cooperatively while it computes the current block and stages them in LDS,
// where every lane reads them by broadcast; the next block's right-hand side
// is prefetched into registers.  Same operations in the same order as
// band_solve_kernel.

// rows of block ez (FIRST: ez == 0, no previous block): y = b - L y
template <int BS, bool FIRST>
__device__ __forceinline__ void band_fwd_rows(const double* __restrict__ blk, double* cur, const double* prev) {
  constexpr int bw = BS - 1, nb = 2 * bw + 1;
#pragma unroll
  for (int m = 0; m < BS; ++m) {
    double v = cur[m];
    if constexpr (!FIRST) {
#pragma unroll
      for (int p = m + 1; p < BS; ++p) v -= blk[m * nb + bw + p - m - BS] * prev[p];
    }
#pragma unroll
    for (int q = 0; q < m; ++q) v -= blk[m * nb + bw + q - m] * cur[q];
    cur[m] = v;
  }
}

// rows of block ez backwards (LAST: ez == KV-1, no next block): x = (y - U x) / U_rr
template <int BS, bool LAST>
__device__ __forceinline__ void band_bwd_rows(const double* __restrict__ blk, double* cur, const double* next) {
  constexpr int bw = BS - 1, nb = 2 * bw + 1;
#pragma unroll
  for (int m = BS - 1; m >= 0; --m) {
    double v = cur[m];
#pragma unroll
    for (int q = m + 1; q < BS; ++q) v -= blk[m * nb + bw + q - m] * cur[q];
    if constexpr (!LAST) {
#pragma unroll
      for (int p = 0; p < m; ++p) v -= blk[m * nb + bw + BS + p - m] * next[p];
    }
    cur[m] = v / blk[m * nb + bw];
  }
}

template <int NS, int NQ>
__global__ void __launch_bounds__(64)
band_solve_reg_kernel(BandGeo g, const double* __restrict__ fac, const int32_t* __restrict__ perm,
                      const int32_t* __restrict__ wave_mat, double* x, const double* b) {
  constexpr int BS = NS * NQ, bw = BS - 1, nb = 2 * bw + 1;
  constexpr int FB = BS * nb;          // factor entries of one block of rows
  constexpr int FPL = (FB + 63) / 64;  // per lane
  __shared__ double fbuf[2][FB];
  const int lane = threadIdx.x;
  const int32_t pl = perm[blockIdx.x * (int64_t)64 + lane];
  const bool active = pl >= 0;
  const int64_t line = active ? pl : 0;  // padding lanes read line 0 and store nothing
  const double* F = fac + (int64_t)wave_mat[blockIdx.x] * ((int64_t)g.n * nb);
  const int KV = g.KV;
  const int64_t c = line / g.Nh;
  const int h = (int)(line - c * g.Nh);
  const int32_t* col = g.colelem + c * KV;
  auto gi = [&](int ez, int m) -> int64_t {
    return (int64_t)(m % NS) * g.Ntot + h + (int64_t)g.Np * col[ez] + (int64_t)g.Nh * (m / NS);
  };
  double pf[FPL];
  auto fetch = [&](int ez) {
#pragma unroll
    for (int i = 0; i < FPL; ++i) {
      const int k = lane + 64 * i;
      if (k < FB) pf[i] = F[(int64_t)ez * FB + k];
    }
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int i = 0; i < FPL; ++i) {
      const int k = lane + 64 * i;
      if (k < FB) fbuf[buf][k] = pf[i];
    }
  };
  auto load = [&](const double* src, int ez, double* dst) {
#pragma unroll
    for (int m = 0; m < BS; ++m) dst[m] = src[gi(ez, m)];
  };
  auto store = [&](int ez, const double* v) {
    if (active) {
#pragma unroll
      for (int m = 0; m < BS; ++m) x[gi(ez, m)] = v[m];
    }
  };
  double prev[BS], cur[BS], nxt[BS];
  // ---- forward: L y = b (y stored in x)
  fetch(0);
  stage(0);
  load(b, 0, nxt);
  __syncthreads();
  for (int ez = 0; ez < KV; ++ez) {
    const int buf = ez & 1;
#pragma unroll
    for (int m = 0; m < BS; ++m) cur[m] = nxt[m];
    if (ez + 1 < KV) {
      fetch(ez + 1);
      load(b, ez + 1, nxt);
    }
    if (ez == 0) band_fwd_rows<BS, true>(fbuf[buf], cur, prev);
    else band_fwd_rows<BS, false>(fbuf[buf], cur, prev);
    store(ez, cur);
#pragma unroll
    for (int m = 0; m < BS; ++m) prev[m] = cur[m];
    if (ez + 1 < KV) stage(1 - buf);
    __syncthreads();
  }
  // ---- backward: U x = y
  fetch(KV - 1);
  stage(0);
  load(x, KV - 1, nxt);
  __syncthreads();
  for (int ez = KV - 1, it = 0; ez >= 0; --ez, ++it) {
    const int buf = it & 1;
#pragma unroll
    for (int m = 0; m < BS; ++m) cur[m] = nxt[m];
    if (ez > 0) {
      fetch(ez - 1);
      load(x, ez - 1, nxt);
    }
    if (ez == KV - 1) band_bwd_rows<BS, true>(fbuf[buf], cur, prev);
    else band_bwd_rows<BS, false>(fbuf[buf], cur, prev);
    store(ez, cur);
#pragma unroll
    for (int m = 0; m < BS; ++m) prev[m] = cur[m];
    if (ez > 0) stage(1 - buf);
    __syncthreads();
  }
}

// Block form of one factored matrix u, element block i (one workgroup per (u, i)):
// the coupling blocks L_{i,i-1} (entries with p > m) and U_{i,i+1} (p < m), and the
// inverses of the unit lower L_ii and upper U_ii by substitution on identity
// columns (one thread per column).
__global__ void band_blocks_kernel(BandGeo g, const double* __restrict__ fac, double* __restrict__ blocks) {
  const int64_t ui = blockIdx.x;
  const int64_t u = ui / g.KV;
  const int i = (int)(ui - u * g.KV);
  const int BS = g.bw + 1, nb = g.nb, bw = g.bw, BB = BS * BS;
  const double* F = fac + u * ((int64_t)g.n * nb);
  double* B = blocks + ui * 4 * (int64_t)BB;
  auto A = [&](int r, int c) -> double { return F[(int64_t)r * nb + bw + c - r]; };
  for (int t = threadIdx.x; t < BB; t += blockDim.x) {
    const int m = t / BS, p = t - m * BS, r = i * BS + m;
    B[t] = (i > 0 && p > m) ? A(r, (i - 1) * BS + p) : 0.0;
    B[2 * BB + t] = (i + 1 < g.KV && p < m) ? A(r, (i + 1) * BS + p) : 0.0;
  }
  const int c = threadIdx.x;
  if (c < BS) {
    double* Li = B + BB;
    double* Ui = B + 3 * BB;
    for (int m = 0; m < BS; ++m) {
      double v = (m == c) ? 1.0 : 0.0;
      for (int k = c; k < m; ++k) v -= A(i * BS + m, i * BS + k) * Li[k * BS + c];
      Li[m * BS + c] = v;
    }
    for (int m = BS - 1; m >= 0; --m) {
      double v = (m == c) ? 1.0 : 0.0;
      for (int k = m + 1; k <= c; ++k) v -= A(i * BS + m, i * BS + k) * Ui[k * BS + c];
      Ui[m * BS + c] = v / A(i * BS + m, i * BS + m);
    }
  }
}

// Banded solve in block form: per element block of rows, forward
//   y_i = L_ii^-1 (b_i - L_{i,i-1} y_{i-1}),   backward  x_i = U_ii^-1 (y_i - U_{i,i+1} x_{i+1}),
// two dense BS x BS products per block with no serial chain.  A wave takes 16
// lines (grouped by matrix, so one matrix per wave), 4 lanes per line, each
// lane forming ceil(BS/4) rows of every product; the vector of a line is
// exchanged through LDS between the products.  The next block's matrices are
// loaded while the current block is computed and staged in LDS (read by
// broadcast).  Agrees with substitution to round-off.
template <int NS, int NQ>
__global__ void __launch_bounds__(64)
band_solve_blk_kernel(BandGeo g, const double* __restrict__ blocks, const int32_t* __restrict__ perm,
                      const int32_t* __restrict__ wave_mat, double* x, const double* b) {
#pragma clang fp contract(fast)
  constexpr int BS = NS * NQ, PARTS = 4, LPW = 64 / PARTS, RPL = (BS + PARTS - 1) / PARTS;
  constexpr int BB = BS * BS, FB = 2 * BB, FPL = (FB + 63) / 64;
  __shared__ double mats[2][FB];
  __shared__ double vec[LPW][BS + 1];
  const int lane = threadIdx.x, j = lane / PARTS, part = lane % PARTS;
  const int64_t line = perm[blockIdx.x * (int64_t)LPW + j];  // padding lines repeat a line of the run
  const int KV = g.KV;
  const double* Bu = blocks + (int64_t)wave_mat[blockIdx.x] * KV * 2 * FB;  // block i: [fwd FB | bwd FB]
  const int64_t c = line / g.Nh;
  const int h = (int)(line - c * g.Nh);
  const int32_t* col = g.colelem + c * KV;
  auto gi = [&](int ez, int m) -> int64_t {
    return (int64_t)(m % NS) * g.Ntot + h + (int64_t)g.Np * col[ez] + (int64_t)g.Nh * (m / NS);
  };
  double pf[FPL];
  auto fetch = [&](int ez, int half) {
#pragma unroll
    for (int i = 0; i < FPL; ++i) {
      const int k = lane + 64 * i;
      if (k < FB) pf[i] = Bu[((int64_t)ez * 2 + half) * FB + k];
    }
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int i = 0; i < FPL; ++i) {
      const int k = lane + 64 * i;
      if (k < FB) mats[buf][k] = pf[i];
    }
  };
  auto load_rows = [&](const double* src, int ez, double* dst) {
#pragma unroll
    for (int k = 0; k < RPL; ++k) {
      const int m = part * RPL + k;
      dst[k] = (m < BS) ? src[gi(ez, m)] : 0.0;
    }
  };
  double v[BS], rb[RPL], nrb[RPL], t[RPL];
  for (int pass = 0; pass < 2; ++pass) {
    const double* src = pass == 0 ? b : x;
    const int first = pass == 0 ? 0 : KV - 1, step = pass == 0 ? 1 : -1;
    fetch(first, pass);
    stage(0);
    load_rows(src, first, nrb);
    __syncthreads();
    for (int it = 0, ez = first; it < KV; ++it, ez += step) {
      const int buf = it & 1;
#pragma unroll
      for (int k = 0; k < RPL; ++k) rb[k] = nrb[k];
      if (it + 1 < KV) {
        fetch(ez + step, pass);
        load_rows(src, ez + step, nrb);
      }
      const double* Cm = mats[buf];       // coupling block
      const double* Iv = mats[buf] + BB;  // inverse of the diagonal block
      // t = r - C v (v: the neighbouring block's solution); rows past BS (when 4 does
      // not divide BS) compute on row BS-1 and are never stored
#pragma unroll
      for (int k = 0; k < RPL; ++k) {
        const int m = part * RPL + k < BS ? part * RPL + k : BS - 1;
        double s = rb[k];
        if (it > 0) {
#pragma unroll
          for (int p = 0; p < BS; ++p) s -= Cm[m * BS + p] * v[p];
        }
        t[k] = s;
      }
#pragma unroll
      for (int k = 0; k < RPL; ++k) {
        const int m = part * RPL + k;
        if (m < BS) vec[j][m] = t[k];
      }
      __syncthreads();
#pragma unroll
      for (int p = 0; p < BS; ++p) v[p] = vec[j][p];
      // y = D^-1 t
#pragma unroll
      for (int k = 0; k < RPL; ++k) {
        const int m = part * RPL + k < BS ? part * RPL + k : BS - 1;
        double s = 0.0;
#pragma unroll
        for (int p = 0; p < BS; ++p) s += Iv[m * BS + p] * v[p];
        t[k] = s;
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < RPL; ++k) {
        const int m = part * RPL + k;
        if (m < BS) {
          vec[j][m] = t[k];
          x[gi(ez, m)] = t[k];
        }
      }
      if (it + 1 < KV) stage(1 - buf);
      __syncthreads();
#pragma unroll
      for (int p = 0; p < BS; ++p) v[p] = vec[j][p];
    }
    __syncthreads();
  }
}

// Fragments of the block form for v_mfma_f64_4x4x4_4b_f64, per matrix u, element
// block i and matrix mi (0: -L_{i,i-1}, 1: L_ii^-1, 2: -U_{i,i+1}, 3: U_ii^-1), in
// 4 x 4 tiles (ti, tk): 16 doubles, element (r, c) of the tile at r + 4 c, i.e.
// M[4 ti + r][4 tk + c] (zero outside BS x BS).  Lane l takes element
// (l & 3, l >> 4) as its A operand, the same for the instruction's four blocks.
__global__ void band_frags_kernel(BandGeo g, const double* __restrict__ blocks, double* __restrict__ frags, int KT) {
  const int64_t ui = blockIdx.x;
  const int BS = g.bw + 1, BB = BS * BS, NT = KT * KT;
  const double* B = blocks + ui * 4 * (int64_t)BB;
  double* F = frags + ui * 4 * (int64_t)NT * 16;
  for (int t = threadIdx.x; t < 4 * NT * 16; t += blockDim.x) {
    const int el = t % 16, tile = (t / 16) % NT, mi = t / (16 * NT);
    const int row = 4 * (tile / KT) + (el & 3), col = 4 * (tile % KT) + (el >> 2);
    const double v = (row < BS && col < BS) ? B[mi * BB + row * BS + col] : 0.0;
    F[t] = (mi == 0 || mi == 2) ? -v : v;
  }
}

// Banded solve in block form on the matrix cores with v_mfma_f64_4x4x4_4b_f64.  A
// workgroup of W waves takes 16 W lines of one matrix; each wave takes 16 of them and
// the instruction's four blocks are four groups of four lines (line lw = lane & 15 of
// the wave).  For block b and lanes l = j + 4 b + 16 q the instruction takes A[i][k]
// from lane i + 4 b + 16 k, B[k][j] from lane j + 4 b + 16 k and returns D[i][j] in
// lane j + 4 b + 16 i, so the line vector is held one double per 4-row tile (lane: row
// 4 tk + (lane >> 4) of line lw) and a product's result tile ti is at once the B
// operand of tile tk = ti of the next product.  Per block of rows
//   t = r - C v  (C negated in the fragments, r as accumulator),  v = D^-1 t,
// one instruction per nonzero 4 x 4 tile (the coupling blocks are strictly
// triangular and the diagonal inverses triangular, so about half the tiles are
// skipped), the row tiles of a product interleaved as independent accumulator
// chains.  The 4 x 4 tiles waste no matrix-core work on padding rows and each
// instruction occupies the matrix core 16 cycles.
//
// The operator fragments of a block (the two matrices of the pass, 2 NT tiles of 16
// doubles) are loaded once per workgroup, each thread a few 16-byte pieces, two
// blocks ahead into registers, written to one of two LDS buffers a block ahead and
// read from there by every wave (four lanes share each element).  The right-hand-side
// rows of the forward pass are loaded two blocks ahead into registers; the forward
// solution is kept in LDS (yb[r * (16 W + 1) + line]); the backward pass reads its
// rows from there and stores each finished block at once.  The block loops run whole
// rounds of three blocks with the computation of blocks past the column's end
// skipped (the look-ahead loads are clamped to the last block and issued
// unconditionally, and no loop exit sits between a load and its use): every register
// buffer stays in place across iterations, so the compiler never waits for all
// outstanding loads.  Padding lanes of a run repeat the run's first line: they
// compute the same values and store them to the same place, so every lane stores.
// With NK > 0 the right-hand side is not read from b but formed in the row loads from
// NK terms of an ARK stage combination (ArkRhs).
template <int NS, int NQ, int W>
struct BandMfmaShape {
  static constexpr int BS = NS * NQ, KT = (BS + 3) / 4, NT = KT * KT, LDY = 16 * W + 1;
  static constexpr int FSET = 2 * NT * 16;                                           // doubles of a pass's fragments
  static constexpr int FPAD = (FSET + 128 * W - 1) / (128 * W) * (128 * W), CH = FPAD / (128 * W);  // 16-byte pieces per thread
};

typedef double atm_v2d __attribute__((ext_vector_type(2)));

template <int NS, int NQ, int W, int NK>
__global__ void __launch_bounds__(64 * W)
band_solve_mfma_kernel(BandGeo g, const double* __restrict__ frags, const int32_t* __restrict__ perm,
                       const int32_t* __restrict__ run_mat, double* x, const double* b, ArkRhs ark) {
  using S = BandMfmaShape<NS, NQ, W>;
  constexpr int BS = S::BS, KT = S::KT, NT = S::NT, LDY = S::LDY, FPAD = S::FPAD, CH = S::CH;
  constexpr bool FULL = BS % 4 == 0;  // every lane's rows 4 tk + q lie inside the block
  constexpr int LA = 2, NBUF = LA + 1;
  extern __shared__ double lds[];
  const int KV = g.KV, n = g.n;
  double* yb = lds;                                                 // n x LDY
  double* fbuf = lds + ((size_t)n * LDY + 1) / 2 * 2;               // 2 x FPAD, 16-byte aligned
  int32_t* cl = reinterpret_cast<int32_t*>(fbuf + 2 * FPAD);        // 16 W x KV
  const int tid = threadIdx.x, l = tid & 63, lw = l & 15, q = l >> 4, fe = (l & 3) + 4 * q;
  const int lg = 16 * (tid >> 6) + lw;  // line of the lane in the workgroup
  const int64_t line = perm[blockIdx.x * (int64_t)(16 * W) + lg];
  const double* Fu = frags + (int64_t)run_mat[blockIdx.x] * KV * 4 * NT * 16;
  const int64_t c = line / g.Nh;
  const int h = (int)(line - c * g.Nh);
  for (int ez = q; ez < KV; ez += 4) cl[lg * KV + ez] = g.colelem[c * KV + ez];
  auto gaddr = [&](int ez, int m) -> int64_t {
    return (int64_t)(m % NS) * g.Ntot + h + (int64_t)g.Nh * (m / NS) + (int64_t)g.Np * cl[lg * KV + ez];
  };
  auto row_ok = [&](int m) -> bool { return FULL || m < BS; };
  // 4 x 4 tiles that hold only exact zeros (their instructions would add +0; skipped with
  // their operand reads): L_{i,i-1} is strictly upper, L_ii^-1 lower, U_{i,i+1} strictly
  // lower, U_ii^-1 upper triangular.
  auto tile_used = [](int mi, int ti, int tk) -> bool {
    const int r0 = 4 * ti, r1 = (4 * ti + 3 < BS - 1) ? 4 * ti + 3 : BS - 1;
    const int c0 = 4 * tk, c1 = (4 * tk + 3 < BS - 1) ? 4 * tk + 3 : BS - 1;
    if (mi == 0) return r0 < c1;   // some r < c
    if (mi == 1) return r1 >= c0;  // some r >= c
    if (mi == 2) return r1 > c0;   // some r > c
    return r0 <= c1;               // some r <= c
  };
  // this thread's pieces of a block's fragments (the allocation is padded by FPAD doubles,
  // so the pieces past the last block's set stay inside it)
  auto fetch_share = [&](int ez, int pass, atm_v2d* dst) {
    const atm_v2d* src = reinterpret_cast<const atm_v2d*>(Fu + (int64_t)(ez * 4 + 2 * pass) * NT * 16);
#pragma unroll
    for (int k = 0; k < CH; ++k) dst[k] = src[tid + 64 * W * k];
  };
  auto put_share = [&](int buf, const atm_v2d* src) {
    atm_v2d* d = reinterpret_cast<atm_v2d*>(fbuf + buf * FPAD);
#pragma unroll
    for (int k = 0; k < CH; ++k) d[tid + 64 * W * k] = src[k];
  };
  auto fetch_rows = [&](int ez, double* dst) {
#pragma unroll
    for (int tk = 0; tk < KT; ++tk) {
      const int m = 4 * tk + q;
      if constexpr (NK == 0) {
        dst[tk] = row_ok(m) ? b[gaddr(ez, m)] : 0.0;
      } else {  // NK terms of an ARK stage combination (ark_combine_kernel's expression)
        double r = 0.0;
        if (row_ok(m)) {
          const int64_t ad = gaddr(ez, m);
          r = ark.q[ad];
#pragma unroll
          for (int k = 0; k < NK; ++k) {
            const double im = ark.i[k][ad], ex = ark.unsplit ? ark.e[k][ad] - im : ark.e[k][ad];
            r = r + ark.dt * (ark.c.ea[k] * ex + ark.c.ia[k] * im);
          }
        }
        dst[tk] = r;
      }
    }
  };
  auto lds_rows = [&](int ez, double* dst) {
#pragma unroll
    for (int tk = 0; tk < KT; ++tk) {
      const int m = 4 * tk + q;
      dst[tk] = row_ok(m) ? yb[(ez * BS + m) * LDY + lg] : 0.0;
    }
  };
  double v[KT];
  // one block of rows: t = r - C v (it > 0), v = D^-1 t; fragments from LDS buffer fa
  auto block = [&](int it, int ez, const double* fa, const double* rr, int pass) {
    double a[2 * NT];
#pragma unroll
    for (int f = 0; f < 2 * NT; ++f)
      if (tile_used(2 * pass + f / NT, (f % NT) / KT, (f % NT) % KT)) a[f] = fa[f * 16 + fe];
    double t[KT];
#pragma unroll
    for (int ti = 0; ti < KT; ++ti) t[ti] = rr[ti];
    if (it > 0) {
#pragma unroll
      for (int tk = 0; tk < KT; ++tk)
#pragma unroll
        for (int ti = 0; ti < KT; ++ti)
          if (tile_used(2 * pass, ti, tk))
            t[ti] = __builtin_amdgcn_mfma_f64_4x4x4f64(a[ti * KT + tk], v[tk], t[ti], 0, 0, 0);
    }
#pragma unroll
    for (int ti = 0; ti < KT; ++ti) v[ti] = 0.0;
#pragma unroll
    for (int tk = 0; tk < KT; ++tk)
#pragma unroll
      for (int ti = 0; ti < KT; ++ti)
        if (tile_used(2 * pass + 1, ti, tk))
          v[ti] = __builtin_amdgcn_mfma_f64_4x4x4f64(a[NT + ti * KT + tk], t[tk], v[ti], 0, 0, 0);
#pragma unroll
    for (int ti = 0; ti < KT; ++ti) {
      const int m = 4 * ti + q;
      if (row_ok(m)) {
        if (pass == 0) yb[(ez * BS + m) * LDY + lg] = v[ti];
        else x[gaddr(ez, m)] = v[ti];
      }
    }
  };
  atm_v2d fr[NBUF][CH];   // fragment pieces of the blocks in flight
  double rbuf[NBUF][KT];  // right-hand-side rows of the forward pass in flight
  auto fwd = [&](int ez) { return ez < KV ? ez : KV - 1; };
  auto bwd = [&](int ez) { return ez >= 0 ? ez : 0; };
#pragma unroll
  for (int tk = 0; tk < KT; ++tk) v[tk] = 0.0;
  __syncthreads();  // cl
  // forward: L y = b.  Round it: blocks it + u (u < NBUF) in register buffers u.  Before
  // block i is computed, the loads of block i + LA are issued into the buffer of block i - 1
  // and the fragments of block i + 1 (loaded a block ago) are written to LDS buffer
  // (i + 1) & 1, which block i - 1 finished reading before the last barrier.
#pragma unroll
  for (int u = 0; u < LA; ++u) {
    fetch_share(fwd(u), 0, fr[u]);
    fetch_rows(fwd(u), rbuf[u]);
  }
  put_share(0, fr[0]);
  __syncthreads();
  for (int it = 0; it < KV; it += NBUF) {
#pragma unroll
    for (int u = 0; u < NBUF; ++u) {
      const int i = it + u;
      fetch_share(fwd(i + LA), 0, fr[(u + LA) % NBUF]);
      fetch_rows(fwd(i + LA), rbuf[(u + LA) % NBUF]);
      put_share((i + 1) & 1, fr[(u + 1) % NBUF]);
      if (i < KV) block(i, i, fbuf + (i & 1) * FPAD, rbuf[u], 0);
      __syncthreads();
    }
  }
  // backward: U x = y, rows from LDS, solution stored per block
#pragma unroll
  for (int u = 0; u < LA; ++u) fetch_share(bwd(KV - 1 - u), 1, fr[u]);
  put_share(0, fr[0]);
  __syncthreads();
  for (int it = 0; it < KV; it += NBUF) {
#pragma unroll
    for (int u = 0; u < NBUF; ++u) {
      const int i = it + u, ez = KV - 1 - i;
      fetch_share(bwd(ez - LA), 1, fr[(u + LA) % NBUF]);
      put_share((i + 1) & 1, fr[(u + 1) % NBUF]);
      if (ez >= 0) {
        lds_rows(ez, rbuf[u]);
        block(i, ez, fbuf + (i & 1) * FPAD, rbuf[u], 1);
      }
      __syncthreads();
    }
  }
}

// waves per workgroup of the matrix-core solve: 4 when the workgroup's LDS fits, else 1
size_t band_mfma_lds_bytes(int n, int KV, int BS, int W) {
  const int KT = (BS + 3) / 4, NT = KT * KT, FSET = 2 * NT * 16;
  const int FPAD = (FSET + 128 * W - 1) / (128 * W) * (128 * W);
  return ((size_t)n * (16 * W + 1) + 1) / 2 * 2 * sizeof(double) + (size_t)2 * FPAD * sizeof(double) +
         (size_t)16 * W * KV * sizeof(int32_t);
}

// doubles of padding after the fragments: a workgroup reads a block's fragments in whole
// rounds of 128 W doubles (FPAD), up to 128 W - 1 past the block's FSET, so the last block's
// rounds stay inside the allocation with 128 x 4 (the largest W) doubles after it
size_t band_mfma_frag_pad() { return (size_t)128 * 4; }

typedef void (*BandSolveMfmaFn)(BandGeo, const double*, const int32_t*, const int32_t*, double*, const double*,
                                ArkRhs);

// the matrix-core solve for (Ns, Nq), W waves per workgroup, with the right-hand side read
// (nk = 0) or formed from one ARK stage term (nk = 1; the implicit laws' Ns = 4, 5 only).  Two
// terms (five vectors read per row) made the solve slower than a combination pass before it.
BandSolveMfmaFn band_mfma_kernel(int Ns, int Nq, int W, int nk = 0) {
#define ATM_BAND_NQ(NS, WW, NKK)                                \
  switch (Nq) {                                                 \
    case 2: return band_solve_mfma_kernel<NS, 2, WW, NKK>;      \
    case 3: return band_solve_mfma_kernel<NS, 3, WW, NKK>;      \
    case 4: return band_solve_mfma_kernel<NS, 4, WW, NKK>;      \
    case 5: return band_solve_mfma_kernel<NS, 5, WW, NKK>;      \
    case 6: return band_solve_mfma_kernel<NS, 6, WW, NKK>;      \
    default: return nullptr;                                    \
  }
#define ATM_BAND_NS(WW, NKK)                  \
  if (Ns == 4) ATM_BAND_NQ(4, WW, NKK)        \
  if (Ns == 5) ATM_BAND_NQ(5, WW, NKK)
  if (nk == 0) {
    if (W == 4) {
      if (Ns == 1) ATM_BAND_NQ(1, 4, 0)
      ATM_BAND_NS(4, 0)
    } else if (W == 1) {
      if (Ns == 1) ATM_BAND_NQ(1, 1, 0)
      ATM_BAND_NS(1, 0)
    }
  } else if (nk == 1) {
    if (W == 4) ATM_BAND_NS(4, 1)
    if (W == 1) ATM_BAND_NS(1, 1)
  }
#undef ATM_BAND_NS
#undef ATM_BAND_NQ
  return nullptr;
}

typedef void (*BandSolveBlkFn)(BandGeo, const double*, const int32_t*, const int32_t*, double*, const double*);

BandSolveBlkFn band_blk_kernel(int Ns, int Nq) {
#define ATM_BAND_NQ(NS)                                \
  switch (Nq) {                                        \
    case 2: return band_solve_blk_kernel<NS, 2>;       \
    case 3: return band_solve_blk_kernel<NS, 3>;       \
    case 4: return band_solve_blk_kernel<NS, 4>;       \
    case 5: return band_solve_blk_kernel<NS, 5>;       \
    case 6: return band_solve_blk_kernel<NS, 6>;       \
    default: return nullptr;                           \
  }
  if (Ns == 1) ATM_BAND_NQ(1)
  if (Ns == 4) ATM_BAND_NQ(4)
  if (Ns == 5) ATM_BAND_NQ(5)
#undef ATM_BAND_NQ
  return nullptr;
}

typedef void (*BandSolveRegFn)(BandGeo, const double*, const int32_t*, const int32_t*, double*, const double*);

// register-blocked kernels for the state counts of the vertical linear laws
// (advection 1, linear Euler-gravity 4 and 5) and NQ = 2..6
BandSolveRegFn band_reg_kernel(int Ns, int Nq) {
#define ATM_BAND_NQ(NS)                                \
  switch (Nq) {                                        \
    case 2: return band_solve_reg_kernel<NS, 2>;       \
    case 3: return band_solve_reg_kernel<NS, 3>;       \
    case 4: return band_solve_reg_kernel<NS, 4>;       \
    case 5: return band_solve_reg_kernel<NS, 5>;       \
    case 6: return band_solve_reg_kernel<NS, 6>;       \
    default: return nullptr;                           \
  }
  if (Ns == 1) ATM_BAND_NQ(1)
  if (Ns == 4) ATM_BAND_NQ(4)
  if (Ns == 5) ATM_BAND_NQ(5)
#undef ATM_BAND_NQ
  return nullptr;
}

constexpr int BAND_SOLVE_PARTS = 4;

size_t band_solve_lds_bytes(int n) { return (size_t)n * (64 / BAND_SOLVE_PARTS + 1) * sizeof(double); }

BandGeo band_geo(const atm_banded* b) {
  BandGeo g;
  g.Ns = b->lin->Ns;
  g.Nq = b->lin->Nq;
  g.Nh = b->Nh;
  g.Np = b->lin->Np;
  g.KV = b->KV;
  g.n = b->n;
  g.bw = b->bw;
  g.nb = b->nb;
  g.nlines = b->nlines;
  g.Ntot = b->lin->Ntot;
  g.elem_col = b->elem_col;
  g.elem_ez = b->elem_ez;
  g.colelem = b->colelem;
  return g;
}

// ---- LSRK tables (src/odesolvers.jl:148-229) ---------------------------------
struct LSRKTable {
  int ns;
  double a[14], b[14];
};

LSRKTable lsrk_table(int scheme) {
  LSRKTable t{};
  if (scheme == ATM_LSRK54) {
    const double a[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                         -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
    const double b[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                         1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                         2277821191437.0 / 14882151754819.0};
    t.ns = 5;
    for (int i = 0; i < 5; ++i) { t.a[i] = a[i]; t.b[i] = b[i]; }
  } else {
    const double a[14] = {0, -0.7188012108672410, -0.7785331173421570, -0.0053282796654044,
                          -0.8552979934029281, -3.9564138245774565, -1.5780575380587385,
                          -2.0837094552574054, -0.7483334182761610, -0.7032861106563359,
                          0.0013917096117681, -0.0932075369637460, -0.9514200470875948,
                          -7.1151571693922548};
    const double b[14] = {0.0367762454319673, 0.3136296607553959, 0.1531848691869027,
                          0.0030097086818182, 0.3326293790646110, 0.2440251405350864,
                          0.3718879239592277, 0.6204126221582444, 0.1524043173028741,
                          0.0760894927419266, 0.0077604214040978, 0.0024647284755382,
                          0.0780348340049386, 5.5059777270269628};
    t.ns = 14;
    for (int i = 0; i < 14; ++i) { t.a[i] = a[i]; t.b[i] = b[i]; }
  }
  return t;
}

Resolved resolve(const atm_dg_desc* d, int Nq, int Np) {
  const int f = d->volume_form, vf = d->volume_flux, sf = d->surface_flux;
  if (d->precision == ATM_PRECISION_FAST && f == ATM_FORM_FLUX_DIFFERENCING && vf == ATM_FLUX_EC &&
      d->dim >= 2) {
    Resolved r{nullptr, 0, 0, nullptr, 1};
    if (d->law == ATM_LAW_EULER) r = resolve_fast_euler(d->dim, sf, Nq);
    if (d->law == ATM_LAW_EULER_GRAVITY) r = resolve_fast_euler_gravity(d->dim, sf, Nq);
    if (r.fn) return r;
  }
  // the Nq-specialised kernels are built for the :per_dir order; the other two
  // reference orders run the generic kernels (a surface flux no specialisation
  // is selected by; the kernels read the real one from the operator)
  const int sel = d->fd_kernel != ATM_FD_PER_DIR ? -1 : sf;
  switch (d->law) {
    case ATM_LAW_ADVECTION: return resolve_advection(d->dim, f, vf, sel, Nq, Np);
    case ATM_LAW_EULER: return resolve_euler(d->dim, f, vf, sel, Nq, Np);
    case ATM_LAW_EULER_GRAVITY: return resolve_euler_gravity(d->dim, f, vf, sel, Nq, Np);
    case ATM_LAW_LINEAR_EULER_GRAVITY: return resolve_linear_euler_gravity(d->dim, f, vf, sel, Nq, Np);
    case ATM_LAW_SHALLOW_WATER: return resolve_shallow_water(d->dim, f, vf, sel, Nq, Np);
    default: return Resolved{nullptr, 0, 0, nullptr, 1};
  }
}

template <template <int> class LAW>
bool sflux_ok_dims(int dim, int sf) {
  if (dim == 1) return sflux_supported<LAW<1>>(sf);
  if (dim == 2) return sflux_supported<LAW<2>>(sf);
  return sflux_supported<LAW<3>>(sf);
}

bool sflux_ok(int law, int dim, int sf) {
  switch (law) {
    case ATM_LAW_ADVECTION: return sflux_ok_dims<Advection>(dim, sf);
    case ATM_LAW_EULER: return sflux_ok_dims<Euler>(dim, sf);
    case ATM_LAW_EULER_GRAVITY: return sflux_ok_dims<EulerGravity>(dim, sf);
    case ATM_LAW_LINEAR_EULER_GRAVITY: return sflux_ok_dims<LinearEulerGravity>(dim, sf);
    case ATM_LAW_SHALLOW_WATER: return sflux_ok_dims<ShallowWater>(dim, sf);
    default: return false;
  }
}


// reduction buffer of an operator: partials, then RED_STATE relaxation slots, then RED_OUT results
constexpr int64_t RED_PART = (int64_t)RED_MAXBLK * RED_MAXQ * 2;
constexpr int64_t RED_STATE = 16;
constexpr int64_t RED_OUT = 16;

int red_buffer(atm_dg* dg) {
  if (dg->red) return ATM_OK;
  void* p = nullptr;
  HIPCHK(dg->ctx, hipMalloc(&p, (RED_PART + RED_STATE + RED_OUT) * sizeof(double)));
  dg->allocs.push_back(p);
  dg->dev_bytes += (RED_PART + RED_STATE + RED_OUT) * sizeof(double);
  dg->red = static_cast<double*>(p);
  return ATM_OK;
}

double* red_state(atm_dg* dg) { return dg->red + RED_PART; }
double* red_out(atm_dg* dg) { return dg->red + RED_PART + RED_STATE; }

// reduce (mode) then combine into out[0..nq) and copy to the host
int reduce_to_host(atm_dg* dg, int mode, const double* q, const double* k, double* host) {
  atm_ctx* ctx = dg->ctx;
  if (!dg->mfn.reduce) return fail(ctx, ATM_ERR_UNSUPPORTED, "the balance law defines no entropy");
  int rc = red_buffer(dg);
  if (rc) return rc;
  hipStream_t s = stream_of(ctx);
  HIPCHK(ctx, dg->mfn.reduce(dg->dev, s, mode, q, k, nullptr, red_state(dg), dg->red));
  HIPCHK(ctx, launch_finalize(s, dg->red, red_blocks(dg->Ntot), 1, FIN_OUT, 0.0, red_state(dg), red_out(dg)));
  HIPCHK(ctx, hipMemcpyAsync(host, red_out(dg), sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(ctx, hipStreamSynchronize(s));
  return ATM_OK;
}

int launch(atm_dg* dg, const double* q, double* dq, double alpha, int assign, double* qout, double bdt) {
  hipError_t e = dg->fn(dg->dev, stream_of(dg->ctx), q, dq, alpha, assign, qout, bdt);
  if (e != hipSuccess) return hipfail(dg->ctx, e, "element kernel launch");
  return ATM_OK;
}

// ---- fixed scratch of the no-allocation LSRK entry point ----------------------
// The ping-pong state of atm_lsrk_run_noalloc lives in this file-scope device
// array: every device the library runs on holds its own copy from module load
// until the process ends.  In host code the array's name is a host-side stand-in;
// host code uses only the device address hipGetSymbolAddress returns for it on
// each device (resolved once per device when a context is created there).
constexpr int64_t NOALLOC_MAX_DOUBLES = (int64_t)8 << 20;  // 64 MiB
static __device__ __align__(256) double scratch[NOALLOC_MAX_DOUBLES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex noalloc_mu;
bool noalloc_resolved[NOALLOC_MAX_DEVICES] = {};
double* noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

// with `device` current: its copy's address (null if it cannot be resolved)
double* noalloc_scratch_of(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lk(noalloc_mu);
  if (!noalloc_resolved[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    noalloc_ptr[device] = (e == hipSuccess) ? static_cast<double*>(p) : nullptr;
    if (e != hipSuccess) (void)hipGetLastError();
    noalloc_resolved[device] = true;
  }
  return noalloc_ptr[device];
}

}  // namespace

// =============================================================================
extern "C" {

const char* atm_version(void) { return "atum-hip 0.1 gfx950"; }

int atm_ctx_create(int device, atm_ctx** out) {
  if (!out) return ATM_ERR_ARG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return ATM_ERR_HIP;
  if (hipSetDevice(device) != hipSuccess) return ATM_ERR_HIP;
  atm_ctx* c = new atm_ctx();
  c->device = device;
  if (hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return ATM_ERR_HIP;
  }
  c->stream = c->own;
  c->noalloc_scratch = noalloc_scratch_of(device);
  int lds = 0;
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lds > 0)
    c->max_lds = (size_t)lds;
  *out = c;
  return ATM_OK;
}

int atm_ctx_destroy(atm_ctx* ctx) {
  if (!ctx) return ATM_ERR_ARG;
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  if (ctx->copy) hipStreamSynchronize(ctx->copy);
  for (hipEvent_t ev : ctx->events)
    if (ev) hipEventDestroy(ev);
  for (int i = 0; i < 16; ++i) {
    if (ctx->snap_ready[i]) hipEventDestroy(ctx->snap_ready[i]);
    if (ctx->snap_done[i]) hipEventDestroy(ctx->snap_done[i]);
  }
  if (ctx->copy) hipStreamDestroy(ctx->copy);
  if (ctx->own) hipStreamDestroy(ctx->own);
  delete ctx;
  return ATM_OK;
}

int atm_stream_create(atm_ctx* ctx, void** stream) {
  if (!ctx || !stream) return ATM_ERR_ARG;
  hipSetDevice(ctx->device);
  hipStream_t s = nullptr;
  HIPCHK(ctx, hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = (void*)s;
  return ATM_OK;
}

int atm_stream_destroy(atm_ctx* ctx, void* stream) {
  if (!ctx || !stream) return ATM_ERR_ARG;
  hipSetDevice(ctx->device);
  HIPCHK(ctx, hipStreamSynchronize((hipStream_t)stream));
  HIPCHK(ctx, hipStreamDestroy((hipStream_t)stream));
  return ATM_OK;
}

int atm_stream_sync(atm_ctx* ctx, void* stream) {
  if (!ctx || !stream) return ATM_ERR_ARG;
  HIPCHK(ctx, hipStreamSynchronize((hipStream_t)stream));
  return ATM_OK;
}

int atm_stream_wait(atm_ctx* ctx, void* waiter, void* waitee) {
  if (!ctx) return ATM_ERR_ARG;
  hipStream_t a = waiter ? (hipStream_t)waiter : stream_of(ctx);
  hipStream_t b = waitee ? (hipStream_t)waitee : stream_of(ctx);
  if (a == b) return ATM_OK;
  hipSetDevice(ctx->device);
  hipEvent_t ev = nullptr;
  HIPCHK(ctx, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  hipError_t e = hipEventRecord(ev, b);
  if (e == hipSuccess) e = hipStreamWaitEvent(a, ev, 0);
  hipEventDestroy(ev);  // released once the wait has been satisfied
  if (e != hipSuccess) return hipfail(ctx, e, "atm_stream_wait");
  return ATM_OK;
}

int atm_ctx_set_stream(atm_ctx* ctx, void* hip_stream) {
  if (!ctx) return ATM_ERR_ARG;
  ctx->stream = hip_stream ? static_cast<hipStream_t>(hip_stream) : ctx->own;
  return ATM_OK;
}

void* atm_ctx_get_stream(atm_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

const char* atm_last_error(atm_ctx* ctx) {
  // the calling thread's copy stays valid until its next call here
  thread_local std::string copy;
  if (!ctx) return "null context";
  copy = last_error(ctx);
  return copy.c_str();
}

int atm_sync(atm_ctx* ctx) {
  if (!ctx) return ATM_ERR_ARG;
  HIPCHK(ctx, hipStreamSynchronize(stream_of(ctx)));
  return ATM_OK;
}

int atm_alloc(atm_ctx* ctx, size_t bytes, void** dptr) {
  if (!ctx || !dptr) return ATM_ERR_ARG;
  hipSetDevice(ctx->device);
  hipError_t e = hipMalloc(dptr, bytes ? bytes : 8);
  if (e == hipErrorOutOfMemory) return fail(ctx, ATM_ERR_NOMEM, "atm_alloc: out of device memory");
  if (e != hipSuccess) return hipfail(ctx, e, "hipMalloc");
  return ATM_OK;
}

int atm_free(atm_ctx* ctx, void* dptr) {
  if (!ctx) return ATM_ERR_ARG;
  if (dptr) HIPCHK(ctx, hipFree(dptr));
  return ATM_OK;
}

int atm_upload(atm_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx) return ATM_ERR_ARG;
  HIPCHK(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream_of(ctx)));
  HIPCHK(ctx, hipStreamSynchronize(stream_of(ctx)));
  return ATM_OK;
}

int atm_download(atm_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx) return ATM_ERR_ARG;
  HIPCHK(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream_of(ctx)));
  HIPCHK(ctx, hipStreamSynchronize(stream_of(ctx)));
  return ATM_OK;
}

int atm_copy(atm_ctx* ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx) return ATM_ERR_ARG;
  HIPCHK(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream_of(ctx)));
  return ATM_OK;
}

}  // extern "C"

namespace atm {

int dg_alloc(atm_dg* dg, size_t bytes, void** out) {
  *out = nullptr;
  HIPCHK(dg->ctx, hipMalloc(out, bytes));
  dg->allocs.push_back(*out);
  dg->dev_bytes += bytes;
  return ATM_OK;
}

int dg_prepare(atm_ctx* ctx, const atm_dg_desc* d, atm_dg** out) {
  if (!ctx || !d || !out) return fail(ctx, ATM_ERR_ARG, "atm_dg_create: null argument");
  *out = nullptr;
  if (d->dim < 1 || d->dim > 3) return fail(ctx, ATM_ERR_ARG, "dim must be 1, 2 or 3");
  if (d->N < 1 || d->N > 7) return fail(ctx, ATM_ERR_UNSUPPORTED, "polynomial degree must be 1..7");
  if (d->K < 0) return fail(ctx, ATM_ERR_ARG, "K must be >= 0");
  if (d->dirmask <= 0 || d->dirmask >= (1 << d->dim))
    return fail(ctx, ATM_ERR_ARG, "directions must be a non-empty subset of 1:dim");
  if (d->volume_form != ATM_FORM_WEAK && d->volume_form != ATM_FORM_FLUX_DIFFERENCING)
    return fail(ctx, ATM_ERR_ARG, "unknown volume form");
  if (d->precision != ATM_PRECISION_REFERENCE && d->precision != ATM_PRECISION_FAST)
    return fail(ctx, ATM_ERR_ARG, "unknown precision mode");
  if (d->fd_kernel < ATM_FD_PER_DIR || d->fd_kernel > ATM_FD_NAIVE)
    return fail(ctx, ATM_ERR_ARG, "unknown flux-differencing kernel type");
  if (d->fd_kernel != ATM_FD_PER_DIR && d->volume_form != ATM_FORM_FLUX_DIFFERENCING)
    return fail(ctx, ATM_ERR_ARG, "a kernel type applies to the flux-differencing form only");
  if (d->fd_kernel != ATM_FD_PER_DIR && d->precision == ATM_PRECISION_FAST)
    return fail(ctx, ATM_ERR_ARG, "the kernel type selects a reference summation order; FAST precision has its own");
  if (d->fd_kernel == ATM_FD_NAIVE && d->dirmask != (1 << d->dim) - 1)
    return fail(ctx, ATM_ERR_ARG, "the naive flux-differencing kernel needs all directions (src/dgsem.jl:139)");
  const int Nq = d->N + 1;
  int Np = 1;
  for (int i = 0; i < d->dim; ++i) Np *= Nq;
  Resolved r = resolve(d, Nq, Np);
  if (!r.fn) return fail(ctx, ATM_ERR_UNSUPPORTED, "balance law does not define this volume flux");
  if (!sflux_ok(d->law, d->dim, d->surface_flux))
    return fail(ctx, ATM_ERR_UNSUPPORTED, "balance law does not define this surface flux");
  if ((d->law == ATM_LAW_EULER_GRAVITY && d->naux < 1 + d->dim) ||
      (d->law == ATM_LAW_EULER_GRAVITY && d->pde_level_balance && d->naux < 3 + d->dim) ||
      (d->law == ATM_LAW_LINEAR_EULER_GRAVITY && d->naux < 3 + d->dim))
    return fail(ctx, ATM_ERR_ARG, "aux state has too few components for the law");
  if (d->naux > 3 + d->dim) return fail(ctx, ATM_ERR_ARG, "aux state has too many components");
  if (d->naux > 0 && d->K > 0 && !d->aux) return fail(ctx, ATM_ERR_ARG, "missing aux array");
  if (d->bc == ATM_BC_PRESCRIBED && (d->n_bc_states < 1 || !d->bc_states))
    return fail(ctx, ATM_ERR_ARG, "prescribed boundary needs bc_states");
  if (r.lds > ctx->max_lds) return fail(ctx, ATM_ERR_UNSUPPORTED, "element does not fit in LDS");
  const int Nfpf = Np / Nq, Nfp = 2 * d->dim * Nfpf;
  const int64_t Ntot = (int64_t)Np * d->K, NF = (int64_t)Nfp * d->K;
  hipSetDevice(ctx->device);
  atm_dg* dg = new atm_dg();
  dg->ctx = ctx;
  dg->law = d->law;
  dg->dim = d->dim;
  dg->Nq = Nq;
  dg->Np = Np;
  dg->Ns = r.ns;
  dg->Ntot = Ntot;
  dg->fn = r.fn;
  dg->bfn = r.bfn;
  dg->lds = r.lds;
  dg->epw = r.epw;
  dg->lds_fixed = r.lds_fixed;
  dg->lds_elem = r.lds_elem;
  dg->max_threads = r.max_threads;
  dg->mbfn = r.mbfn;
  dg->mfn = resolve_measures(d->law, d->dim);
  DevDG& p = dg->dev;
  p.epw = r.epw;
  p.lds = (int)r.lds;
  p.c.gamma = d->gamma;
  p.c.grav = d->grav;
  p.c.omega = d->omega;
  for (int i = 0; i < 3; ++i) p.c.u[i] = d->u[i];
  p.c.pde_level_balance = d->pde_level_balance;
  p.c.bc = d->bc;
  p.c.source = d->source;
  p.c.n_bc_states = d->n_bc_states;
  p.K = d->K;
  p.Nq = Nq;
  p.Np = Np;
  p.Nfpf = Nfpf;
  p.Nfp = Nfp;
  p.naux = d->naux;
  p.dirmask = d->dirmask;
  p.vflux = d->volume_flux;
  p.sflux = d->surface_flux;
  p.fdk = d->fd_kernel;
  p.Ntot = Ntot;
  p.NF = NF;
  int rc = ATM_OK;
  if ((rc = upload_copy(dg, d->aux, (size_t)d->naux * Ntot, &p.aux)) ||
      (rc = upload_copy(dg, d->bc_states, (size_t)(d->bc == ATM_BC_PRESCRIBED ? d->n_bc_states * r.ns : 0),
                        &p.c.bc_states))) {
    atm_dg_destroy(dg);
    return rc;
  }
  *out = dg;
  return ATM_OK;
}

}  // namespace atm

extern "C" {

int atm_dg_create(atm_ctx* ctx, const atm_dg_desc* d, atm_dg** out) {
  if (!ctx || !d || !out) return fail(ctx, ATM_ERR_ARG, "atm_dg_create: null argument");
  *out = nullptr;
  if (d->dim < 1 || d->dim > 3) return fail(ctx, ATM_ERR_ARG, "dim must be 1, 2 or 3");
  if (d->N < 1 || d->N > 7) return fail(ctx, ATM_ERR_UNSUPPORTED, "polynomial degree must be 1..7");
  if (d->K < 0) return fail(ctx, ATM_ERR_ARG, "K must be >= 0");
  const int Nq = d->N + 1;
  int Np = 1;
  for (int i = 0; i < d->dim; ++i) Np *= Nq;
  const int Nfpf = Np / Nq, Nfp = 2 * d->dim * Nfpf;
  const int64_t Ntot = (int64_t)Np * d->K, NF = (int64_t)Nfp * d->K;
  // Face points must follow the canonical order: face 2r+side, other node
  // indices ascending with the lower dimension fastest.
  if (d->K > 0 && !d->faceix_m) return fail(ctx, ATM_ERR_ARG, "missing faceix_m");
  for (int64_t e = 0; e < d->K; ++e)
    for (int face = 0; face < 2 * d->dim; ++face) {
      const int rr = face / 2, side = face % 2;
      for (int fp = 0; fp < Nfpf; ++fp) {
        int id[3] = {0, 0, 0}, rem = fp;
        for (int dd = 0; dd < d->dim; ++dd) {
          if (dd == rr) continue;
          id[dd] = rem % Nq;
          rem /= Nq;
        }
        id[rr] = side ? Nq - 1 : 0;
        const int64_t want = id[0] + Nq * (id[1] + (int64_t)Nq * id[2]) + (int64_t)Np * e;
        if (d->faceix_m[e * Nfp + face * Nfpf + fp] != want)
          return fail(ctx, ATM_ERR_UNSUPPORTED, "faceix_m is not in canonical face-point order");
      }
    }
  // every index the kernels follow must stay inside the node array / state table
  if (d->K > 0 && (!d->faceix_p || !d->bctag || !d->MJ || !d->MJI || !d->g || !d->D || !d->faceMJ || !d->normal))
    return fail(ctx, ATM_ERR_ARG, "missing geometry array");
  for (int64_t i = 0; i < NF; ++i)
    if (d->faceix_p[i] < 0 || d->faceix_p[i] >= Ntot) return fail(ctx, ATM_ERR_ARG, "faceix_p out of range");
  for (int64_t i = 0; i < (int64_t)2 * d->dim * d->K; ++i)
    if (d->bctag[i] < 0) return fail(ctx, ATM_ERR_ARG, "negative boundary tag");
  atm_dg* dg = nullptr;
  int rc = dg_prepare(ctx, d, &dg);
  if (rc) return rc;
  DevDG& p = dg->dev;
  const size_t dd = (size_t)d->dim * d->dim;
  if ((rc = upload_copy(dg, d->D, (size_t)Nq * Nq, &p.D)) ||
      (rc = upload_copy(dg, d->g, dd * Ntot, &p.g)) ||
      (rc = upload_copy(dg, d->MJ, (size_t)Ntot, &p.MJ)) ||
      (rc = upload_copy(dg, d->MJI, (size_t)Ntot, &p.MJI)) ||
      (rc = upload_copy(dg, d->faceMJ, (size_t)NF, &p.faceMJ)) ||
      (rc = upload_copy(dg, d->normal, (size_t)d->dim * NF, &p.normal)) ||
      (rc = upload_copy(dg, d->faceix_p, (size_t)NF, &p.faceix_p)) ||
      (rc = upload_copy(dg, d->bctag, (size_t)2 * d->dim * d->K, &p.bctag))) {
    atm_dg_destroy(dg);
    return rc;
  }
  *out = dg;
  return ATM_OK;
}

int atm_dg_destroy(atm_dg* dg) {
  if (!dg) return ATM_ERR_ARG;
  hipSetDevice(dg->ctx->device);
  hipStreamSynchronize(stream_of(dg->ctx));
  if (dg->lsrk_exec) hipGraphExecDestroy(dg->lsrk_exec);
  if (dg->sched) hipFree(dg->sched);
  for (void* p : dg->allocs) hipFree(p);
  delete dg;
  return ATM_OK;
}

int atm_dg_nstates(const atm_dg* dg) { return dg ? dg->Ns : ATM_ERR_ARG; }

int atm_dg_get_epw(const atm_dg* dg) { return dg ? dg->epw : ATM_ERR_ARG; }

int atm_dg_set_epw(atm_dg* dg, int epw) {
  if (!dg) return ATM_ERR_ARG;
  const size_t lds = dg->lds_fixed + (size_t)(epw > 0 ? epw : 0) * dg->lds_elem;
  if (epw < 1 || epw > 512 || epw * dg->Np > dg->max_threads || lds > dg->ctx->max_lds)
    return fail(dg->ctx, ATM_ERR_ARG,
                "atm_dg_set_epw: the workgroup would exceed the kernel's thread limit (atm_dg_epw_candidates) or the LDS");
  if (epw == dg->epw) return ATM_OK;
  hipSetDevice(dg->ctx->device);
  HIPCHK(dg->ctx, hipStreamSynchronize(stream_of(dg->ctx)));
  if (dg->lsrk_exec) hipGraphExecDestroy(dg->lsrk_exec);  // captured with the old launch shape
  dg->lsrk_exec = nullptr;
  dg->lsrk_scheme = -1;
  dg->epw = epw;
  dg->lds = lds;
  dg->dev.epw = epw;
  dg->dev.lds = (int)lds;
  return ATM_OK;
}

int atm_dg_set_schedule(atm_dg* dg, const int32_t* order, int64_t n) {
  if (!dg) return ATM_ERR_ARG;
  atm_ctx* ctx = dg->ctx;
  const int64_t K = dg->dev.K;
  if (order && n != K) return fail(ctx, ATM_ERR_ARG, "atm_dg_set_schedule: the order must list every element");
  int32_t* dev = nullptr;
  if (order && K > 0) {
    std::vector<uint8_t> seen((size_t)K, 0);
    for (int64_t i = 0; i < K; ++i) {
      const int64_t v = order[i];
      if (v < 0 || v >= K || seen[(size_t)v])
        return fail(ctx, ATM_ERR_ARG, "atm_dg_set_schedule: the order is not a permutation of the elements");
      seen[(size_t)v] = 1;
    }
    hipSetDevice(ctx->device);
    HIPCHK(ctx, hipMalloc(&dev, (size_t)K * sizeof(int32_t)));
    hipError_t e = hipMemcpy(dev, order, (size_t)K * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      hipFree(dev);
      return hipfail(ctx, e, "atm_dg_set_schedule upload");
    }
  }
  hipSetDevice(ctx->device);
  HIPCHK(ctx, hipStreamSynchronize(stream_of(ctx)));  // launches in flight may read the old order
  if (dg->lsrk_exec) hipGraphExecDestroy(dg->lsrk_exec);  // captured with the old order
  dg->lsrk_exec = nullptr;
  dg->lsrk_scheme = -1;
  if (dg->sched) {
    hipFree(dg->sched);
    dg->dev_bytes -= (size_t)dg->dev.K * sizeof(int32_t);
  }
  dg->sched = dev;
  dg->dev.sched = dev;
  if (dev) dg->dev_bytes += (size_t)dg->dev.K * sizeof(int32_t);
  return ATM_OK;
}

int64_t atm_dg_device_bytes(const atm_dg* dg) { return dg ? (int64_t)dg->dev_bytes : ATM_ERR_ARG; }

int atm_dg_epw_candidates(const atm_dg* dg, int32_t* out, int max) {
  if (!dg || (max > 0 && !out)) return ATM_ERR_ARG;
  // whole waves of 64 lanes: one element per 1, 2, 4 or 8 waves for large elements,
  // 64/128/256/512 lanes of small elements otherwise
  int cand[8], n = 0;
  auto add = [&](int e) {
    if (e < 1 || e * dg->Np > dg->max_threads || dg->lds_fixed + (size_t)e * dg->lds_elem > dg->ctx->max_lds) return;
    for (int i = 0; i < n; ++i)
      if (cand[i] == e) return;
    cand[n++] = e;
  };
  if (dg->Np >= 64) {
    for (int e : {1, 2, 4}) add(e);
  } else {
    for (int t : {64, 128, 256, 512}) add(t / dg->Np);
  }
  for (int i = 0; i < n && i < max; ++i) out[i] = cand[i];
  return n;
}

int atm_dg_autotune(atm_dg* dg, const double* q, double* dq, int reps, int32_t* epw_out, float* us_out, int max) {
  if (!dg || !q || !dq || reps < 1) return fail(dg ? dg->ctx : nullptr, ATM_ERR_ARG, "atm_dg_autotune: bad argument");
  atm_ctx* ctx = dg->ctx;
  int32_t cand[8];
  const int n = atm_dg_epw_candidates(dg, cand, 8);
  if (n < 1) return fail(ctx, ATM_ERR_UNSUPPORTED, "atm_dg_autotune: no workgroup shape fits");
  hipSetDevice(ctx->device);
  hipEvent_t a = nullptr, b = nullptr;
  HIPCHK(ctx, hipEventCreate(&a));
  hipError_t e = hipEventCreate(&b);
  if (e != hipSuccess) {
    hipEventDestroy(a);
    return hipfail(ctx, e, "hipEventCreate");
  }
  const int first = dg->epw;
  int best = first, rc = ATM_OK;
  float best_ms = 0.f, first_ms = -1.f;
  const hipStream_t s = stream_of(ctx);
  for (int i = 0; i < n && rc == ATM_OK; ++i) {
    if ((rc = atm_dg_set_epw(dg, cand[i]))) break;
    for (int w = 0; w < 2 && rc == ATM_OK; ++w) rc = launch(dg, q, dq, 1.0, 1, nullptr, 0.0);
    if (rc) break;
    hipEventRecord(a, s);
    for (int k = 0; k < reps && rc == ATM_OK; ++k) rc = launch(dg, q, dq, 1.0, 1, nullptr, 0.0);
    hipEventRecord(b, s);
    if (rc) break;
    float ms = 0.f;
    if ((e = hipEventSynchronize(b)) != hipSuccess || (e = hipEventElapsedTime(&ms, a, b)) != hipSuccess) {
      rc = hipfail(ctx, e, "atm_dg_autotune timing");
      break;
    }
    if (epw_out && i < max) epw_out[i] = cand[i];
    if (us_out && i < max) us_out[i] = ms * 1e3f / reps;
    if (cand[i] == first) first_ms = ms;
    if (i == 0 || ms < best_ms) {
      best_ms = ms;
      best = cand[i];
    }
  }
  hipEventDestroy(a);
  hipEventDestroy(b);
  // keep the shape the operator had unless another is clearly faster (2 %): near ties
  // are timing noise
  if (rc == ATM_OK && first_ms > 0.f && best_ms > 0.98f * first_ms) best = first;
  const int rc2 = atm_dg_set_epw(dg, rc == ATM_OK ? best : first);
  if (rc) return rc;
  return rc2 ? rc2 : best;
}
int64_t atm_dg_nnodes(const atm_dg* dg) { return dg ? dg->Ntot : ATM_ERR_ARG; }

int atm_dg_tendency(atm_dg* dg, double* dq, const double* q, double time, int increment) {
  (void)time;
  if (!dg || !dq || !q) return fail(dg ? dg->ctx : nullptr, ATM_ERR_ARG, "atm_dg_tendency: null argument");
  hipSetDevice(dg->ctx->device);
  return launch(dg, q, dq, 1.0, increment ? 0 : 1, nullptr, 0.0);
}

int atm_lsrk_stage(atm_dg* dg, const double* q_in, double* q_out, double* dq, double a, double bdt,
                   double time) {
  (void)time;
  if (!dg || !q_in || !q_out || !dq) return fail(dg ? dg->ctx : nullptr, ATM_ERR_ARG, "atm_lsrk_stage: null argument");
  if (q_in == q_out) return fail(dg->ctx, ATM_ERR_ARG, "atm_lsrk_stage: q_out must not alias q_in");
  hipSetDevice(dg->ctx->device);
  return launch(dg, q_in, dq, a, a == 0.0 ? 1 : 0, q_out, bdt);
}

int atm_lsrk_run(atm_dg* dg, int scheme, double* q, double* dq, double* scratch, double dt, double t0,
                 int64_t nsteps, int use_graph) {
  (void)t0;
  if (!dg || !q || !dq || !scratch) return fail(dg ? dg->ctx : nullptr, ATM_ERR_ARG, "atm_lsrk_run: null argument");
  if (scheme != ATM_LSRK54 && scheme != ATM_LSRK144) return fail(dg->ctx, ATM_ERR_ARG, "unknown LSRK scheme");
  if (q == scratch || dq == scratch || q == dq) return fail(dg->ctx, ATM_ERR_ARG, "atm_lsrk_run: buffers alias");
  hipSetDevice(dg->ctx->device);
  const LSRKTable t = lsrk_table(scheme);
  double* bufs[2] = {q, scratch};
  int cur = 0;  // buffer holding the current state
  auto one_step = [&]() -> int {
    for (int st = 0; st < t.ns; ++st) {
      int rc = launch(dg, bufs[cur], dq, t.a[st], t.a[st] == 0.0 ? 1 : 0, bufs[1 - cur], t.b[st] * dt);
      if (rc) return rc;
      cur = 1 - cur;
    }
    return ATM_OK;
  };
  // each step flips the buffer parity when the stage count is odd; a pair of
  // steps always returns the state to q
  const int64_t pair = (t.ns % 2 == 1) ? 2 : 1;
  int64_t done = 0;
  if (use_graph && nsteps >= pair) {
    // reuse the last instantiated graph when the buffers, dt and scheme are the same
    hipStream_t s = stream_of(dg->ctx);
    if (dg->lsrk_exec && (dg->lsrk_scheme != scheme || dg->lsrk_q != q || dg->lsrk_dq != dq ||
                          dg->lsrk_scratch != scratch || dg->lsrk_dt != dt)) {
      HIPCHK(dg->ctx, hipStreamSynchronize(s));
      hipGraphExecDestroy(dg->lsrk_exec);
      dg->lsrk_exec = nullptr;
    }
    if (!dg->lsrk_exec) {
      hipGraph_t graph = nullptr;
      HIPCHK(dg->ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      int rc = ATM_OK;
      for (int64_t i = 0; i < pair && rc == ATM_OK; ++i) rc = one_step();
      hipError_t ec = hipStreamEndCapture(s, &graph);
      if (rc) {
        if (graph) hipGraphDestroy(graph);
        return rc;
      }
      if (ec != hipSuccess) return hipfail(dg->ctx, ec, "hipStreamEndCapture");
      hipGraphExec_t exec = nullptr;
      hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      hipGraphDestroy(graph);
      if (ei != hipSuccess) return hipfail(dg->ctx, ei, "hipGraphInstantiate");
      dg->lsrk_exec = exec;
      dg->lsrk_scheme = scheme;
      dg->lsrk_q = q;
      dg->lsrk_dq = dq;
      dg->lsrk_scratch = scratch;
      dg->lsrk_dt = dt;
    }
    const int64_t reps = nsteps / pair;
    for (int64_t i = 0; i < reps; ++i) HIPCHK(dg->ctx, hipGraphLaunch(dg->lsrk_exec, s));
    done = reps * pair;
  }
  for (; done < nsteps; ++done) {
    int rc = one_step();
    if (rc) return rc;
  }
  if (cur != 0) {
    const int64_t n = (int64_t)dg->Ns * dg->Ntot;
    hipLaunchKernelGGL(copy_kernel, dim3(ew_blocks(n)), dim3(EW_THREADS), 0, stream_of(dg->ctx), q, scratch, n);
    HIPCHK(dg->ctx, hipGetLastError());
  }
  return ATM_OK;
}

int atm_lsrk_run_noalloc(atm_dg* dg, int scheme, double* q, double* dq, double dt, double t0, int64_t nsteps) {
  (void)t0;
  if (!dg || !q || !dq) return fail(dg ? dg->ctx : nullptr, ATM_ERR_ARG, "atm_lsrk_run_noalloc: null argument");
  atm_ctx* ctx = dg->ctx;
  if (scheme != ATM_LSRK54 && scheme != ATM_LSRK144) return fail(ctx, ATM_ERR_ARG, "unknown LSRK scheme");
  if (q == dq) return fail(ctx, ATM_ERR_ARG, "atm_lsrk_run_noalloc: buffers alias");
  if (nsteps < 0) return fail(ctx, ATM_ERR_ARG, "atm_lsrk_run_noalloc: nsteps < 0");
  const int64_t n = (int64_t)dg->Ns * dg->Ntot;
  if (n > NOALLOC_MAX_DOUBLES)
    return fail(ctx, ATM_ERR_ARG,
                "atm_lsrk_run_noalloc: the state has " + std::to_string(n) + " doubles, more than the " +
                    std::to_string(NOALLOC_MAX_DOUBLES) + " of the fixed scratch (use atm_lsrk_run)");
  double* sc = ctx->noalloc_scratch;
  if (!sc) return fail(ctx, ATM_ERR_HIP, "atm_lsrk_run_noalloc: no device address for the fixed scratch");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const LSRKTable t = lsrk_table(scheme);
  double* bufs[2] = {q, sc};
  int cur = 0;  // buffer holding the current state
  for (int64_t step = 0; step < nsteps; ++step)
    for (int st = 0; st < t.ns; ++st) {
      const int rc = launch(dg, bufs[cur], dq, t.a[st], t.a[st] == 0.0 ? 1 : 0, bufs[1 - cur], t.b[st] * dt);
      if (rc) return rc;
      cur = 1 - cur;
    }
  if (cur != 0) {
    hipLaunchKernelGGL(copy_kernel, dim3(ew_blocks(n)), dim3(EW_THREADS), 0, stream_of(ctx), q, sc, n);
    HIPCHK(ctx, hipGetLastError());
  }
  return ATM_OK;
}

int atm_event_record(atm_ctx* ctx, int slot) {
  if (!ctx || slot < 0 || slot >= 16) return ATM_ERR_ARG;
  if (!ctx->events[slot]) HIPCHK(ctx, hipEventCreate(&ctx->events[slot]));
  HIPCHK(ctx, hipEventRecord(ctx->events[slot], stream_of(ctx)));
  return ATM_OK;
}

int atm_event_elapsed_ms(atm_ctx* ctx, int a, int b, float* ms) {
  if (!ctx || a < 0 || a >= 16 || b < 0 || b >= 16 || !ms || !ctx->events[a] || !ctx->events[b]) return ATM_ERR_ARG;
  HIPCHK(ctx, hipEventSynchronize(ctx->events[b]));
  HIPCHK(ctx, hipEventElapsedTime(ms, ctx->events[a], ctx->events[b]));
  return ATM_OK;
}

int atm_axpby(atm_ctx* ctx, double* y, const double* x, double a, double b, int64_t n) {
  if (!ctx || !y || !x) return ATM_ERR_ARG;
  hipLaunchKernelGGL(axpby_kernel, dim3(ew_blocks(n)), dim3(EW_THREADS), 0, stream_of(ctx), y, x, a, b, n);
  HIPCHK(ctx, hipGetLastError());
  return ATM_OK;
}

// ---- measures (src/measures.jl) and relaxation LSRK ----------------------------

int atm_entropy_integral(atm_dg* dg, const double* q, double* out) {
  if (!dg || !q || !out) return fail(dg ? dg->ctx : nullptr, ATM_ERR_ARG, "atm_entropy_integral: null argument");
  hipSetDevice(dg->ctx->device);
  return reduce_to_host(dg, RED_ENTROPY, q, nullptr, out);
}

int atm_entropy_product(atm_dg* dg, const double* p, const double* q, double* out) {
  if (!dg || !p || !q || !out) return fail(dg ? dg->ctx : nullptr, ATM_ERR_ARG, "atm_entropy_product: null argument");
  hipSetDevice(dg->ctx->device);
  return reduce_to_host(dg, RED_PRODUCT, p, q, out);
}

int atm_weightednorm(atm_dg* dg, const double* q, double p, int componentwise, double* out) {
  if (!dg || !q || !out) return fail(dg ? dg->ctx : nullptr, ATM_ERR_ARG, "atm_weightednorm: null argument");
  if (!(p > 0)) return fail(dg->ctx, ATM_ERR_ARG, "atm_weightednorm: p must be positive (or infinity)");
  atm_ctx* ctx = dg->ctx;
  hipSetDevice(ctx->device);
  int rc = red_buffer(dg);
  if (rc) return rc;
  const int Ns = dg->Ns;
  const bool inf = std::isinf(p);
  hipStream_t s = stream_of(ctx);
  HIPCHK(ctx, launch_wnorm(s, q, dg->dev.MJ, Ns, dg->Ntot, p, dg->red));
  HIPCHK(ctx, launch_finalize(s, dg->red, red_blocks(dg->Ntot), Ns, inf ? FIN_MAX : FIN_OUT, 0.0, red_state(dg),
                              red_out(dg)));
  double v[RED_MAXQ];
  HIPCHK(ctx, hipMemcpyAsync(v, red_out(dg), Ns * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(ctx, hipStreamSynchronize(s));
  // per component (sum MJ |f|^p)^(1/p), then the p-norm of the component norms
  for (int c = 0; c < Ns; ++c)
    if (!inf) v[c] = (p == 2.0) ? std::sqrt(v[c]) : ((p == 1.0) ? v[c] : std::pow(v[c], 1.0 / p));
  if (componentwise) {
    for (int c = 0; c < Ns; ++c) out[c] = v[c];
    return ATM_OK;
  }
  double tot = 0.0;
  if (inf) {
    for (int c = 0; c < Ns; ++c) tot = (std::fabs(v[c]) > tot || v[c] != v[c]) ? std::fabs(v[c]) : tot;
  } else if (p == 2.0) {
    for (int c = 0; c < Ns; ++c) tot += v[c] * v[c];
    tot = std::sqrt(tot);
  } else if (p == 1.0) {
    for (int c = 0; c < Ns; ++c) tot += std::fabs(v[c]);
  } else {
    for (int c = 0; c < Ns; ++c) tot += std::pow(std::fabs(v[c]), p);
    tot = std::pow(tot, 1.0 / p);
  }
  out[0] = tot;
  return ATM_OK;
}

int atm_rlsrk_run(atm_dg* dg, int scheme, double* q, double* work, double dt, double t0, int64_t nsteps,
                  double* gamma) {
  (void)t0;
  if (!dg || !q || !work || !gamma) return fail(dg ? dg->ctx : nullptr, ATM_ERR_ARG, "atm_rlsrk_run: null argument");
  if (scheme != ATM_LSRK54 && scheme != ATM_LSRK144) return fail(dg->ctx, ATM_ERR_ARG, "unknown LSRK scheme");
  if (!dg->mfn.reduce) return fail(dg->ctx, ATM_ERR_UNSUPPORTED, "the balance law defines no entropy");
  atm_ctx* ctx = dg->ctx;
  hipSetDevice(ctx->device);
  int rc = red_buffer(dg);
  if (rc) return rc;
  const int64_t n = (int64_t)dg->Ns * dg->Ntot;
  double* dq = work;
  double* q0 = work + n;
  double* k = work + 2 * n;
  if (q >= work && q < work + 3 * n) return fail(ctx, ATM_ERR_ARG, "atm_rlsrk_run: q aliases the work buffer");
  const LSRKTable t = lsrk_table(scheme);
  // b coefficients of the equivalent standard RK method (src/odesolvers.jl:91-95)
  double bfull[14];
  bfull[t.ns - 1] = t.b[t.ns - 1];
  for (int i = t.ns - 2; i >= 0; --i) bfull[i] = t.a[i + 1] * bfull[i + 1] + t.b[i];
  hipStream_t s = stream_of(ctx);
  double* st = red_state(dg);
  const int nblk = red_blocks(dg->Ntot);
  HIPCHK(ctx, launch_set_slot(s, st, RX_GAMMA, *gamma));
  for (int64_t step = 0; step < nsteps; ++step) {
    // q0 = q; eta0 = entropyintegral(q0); deta = -0.0
    hipLaunchKernelGGL(copy_kernel, dim3(ew_blocks(n)), dim3(EW_THREADS), 0, s, q0, q, n);
    HIPCHK(ctx, hipGetLastError());
    HIPCHK(ctx, dg->mfn.reduce(dg->dev, s, RED_ENTROPY, q0, nullptr, nullptr, st, dg->red));
    HIPCHK(ctx, launch_finalize(s, dg->red, nblk, 1, FIN_ETA0, 0.0, st, nullptr));
    for (int i = 0; i < t.ns; ++i) {
      if ((rc = launch(dg, q, k, 1.0, 1, nullptr, 0.0))) return rc;  // k = L(q), increment = false
      HIPCHK(ctx, dg->mfn.stage(dg->dev, s, q, dq, k, t.a[i], t.b[i] * dt, dg->red));
      HIPCHK(ctx, launch_finalize(s, dg->red, nblk, 1, FIN_DETA, bfull[i] * dt, st, nullptr));
    }
    HIPCHK(ctx, launch_diff(s, k, q, q0, n));  // k = q - q0
    // Newton iterations on the device, four per check of the done flag
    for (int chunk = 0; chunk < 13; ++chunk) {
      for (int it = 0; it < 4; ++it) {
        HIPCHK(ctx, dg->mfn.reduce(dg->dev, s, RED_RELAX, nullptr, k, q0, st, dg->red));
        HIPCHK(ctx, launch_finalize(s, dg->red, nblk, 2, FIN_NEWTON, 0.0, st, nullptr));
      }
      double done = 0.0;
      HIPCHK(ctx, hipMemcpyAsync(&done, st + RX_DONE, sizeof(double), hipMemcpyDeviceToHost, s));
      HIPCHK(ctx, hipStreamSynchronize(s));
      if (done != 0.0) break;
    }
    HIPCHK(ctx, launch_relax_apply(s, q, q0, k, st, n));
  }
  HIPCHK(ctx, hipMemcpyAsync(gamma, st + RX_GAMMA, sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(ctx, hipStreamSynchronize(s));
  return ATM_OK;
}

// ---- batched instances --------------------------------------------------------
}  // extern "C"

namespace atm {

constexpr int MULTI_MAX_THREADS = 512;  // launch bound of the multi-degree batched kernels

// (Re)groups a batch's members.  Device and pinned tables stay with the batch:
// group g takes slot g, which is reallocated only when it is too small.  Graphs
// captured with the old groups are dropped (after the stream has drained them).
int batch_set_members(atm_batch* b, int32_t n, atm_dg* const* members) {
  atm_ctx* ctx = b->ctx;
  if (n < 0 || (n > 0 && !members)) return fail(ctx, ATM_ERR_ARG, "atm_batch: bad member list");
  for (int i = 0; i < n; ++i)
    if (!members[i] || members[i]->ctx != ctx || !members[i]->bfn)
      return fail(ctx, ATM_ERR_ARG, "atm_batch: member is null, from another context or not batchable");
  std::vector<atm_batch::Group> groups;
  for (int i = 0; i < n; ++i) {
    atm_dg* dg = members[i];
    int g = -1;
    for (size_t k = 0; k < groups.size(); ++k) {
      const atm_batch::Group& H = groups[k];
      if (dg->mbfn ? (H.multi && H.bfn == dg->mbfn)
                   : (!H.multi && H.bfn == dg->bfn && H.Np == dg->Np && H.epw == dg->epw && H.lds == dg->lds))
        g = (int)k;
    }
    if (g < 0) {
      atm_batch::Group G;
      G.bfn = dg->bfn;
      G.Np = dg->Np;
      G.epw = dg->epw;
      G.threads = ((dg->epw * dg->Np + 63) / 64) * 64;
      G.lds = dg->lds;
      if (dg->mbfn) {
        G.bfn = dg->mbfn;
        G.multi = true;
        G.Np = G.epw = G.threads = 0;
        G.lds = 0;
      }
      groups.push_back(G);
      g = (int)groups.size() - 1;
    }
    atm_batch::Group& G = groups[g];
    G.members.push_back(i);
    if (G.multi) {
      // the multi-degree kernels are built for workgroups of up to 512 threads: a
      // member whose own shape is larger (3-D N = 6 at two elements per workgroup)
      // takes fewer elements per workgroup here
      int e0 = dg->epw;
      while (e0 > 1 && ((e0 * dg->Np + 63) / 64) * 64 > MULTI_MAX_THREADS) --e0;
      G.mepw.push_back(e0);
      G.mlds.push_back((int)(dg->lds_fixed + (size_t)e0 * dg->lds_elem));
      G.threads = std::max(G.threads, ((e0 * dg->Np + 63) / 64) * 64);
      G.lds = std::max(G.lds, dg->lds_fixed + (size_t)e0 * dg->lds_elem);
    }
  }
  // a multi-degree group launches workgroups of its largest member's size: the
  // other members take as many elements per workgroup as fill it (results do not
  // depend on the shape) while their LDS fits
  for (auto& G : groups) {
    if (!G.multi) continue;
    G.lds = 0;
    for (size_t k = 0; k < G.members.size(); ++k) {
      const atm_dg* dg = members[G.members[k]];
      int e = std::max(G.mepw[k], G.threads / dg->Np);
      while (e > G.mepw[k] && dg->lds_fixed + (size_t)e * dg->lds_elem > ctx->max_lds) --e;
      G.mepw[k] = e;
      G.mlds[k] = (int)(dg->lds_fixed + (size_t)e * dg->lds_elem);
      G.lds = std::max(G.lds, (size_t)G.mlds[k]);
    }
  }
  hipSetDevice(ctx->device);
  if (b->exec[0] || b->exec[1]) {
    HIPCHK(ctx, hipStreamSynchronize(stream_of(ctx)));
    for (hipGraphExec_t& x : b->exec)
      if (x) hipGraphExecDestroy(x);
    b->exec[0] = b->exec[1] = nullptr;
  }
  if (b->slots.size() < groups.size()) b->slots.resize(groups.size());
  for (size_t g = 0; g < groups.size(); ++g) {
    atm_batch::Slot& S = b->slots[g];
    groups[g].slot = (int)g;
    const size_t need = groups[g].members.size();
    if (!S.uploaded) {
      hipError_t e = hipEventCreateWithFlags(&S.uploaded, hipEventDisableTiming);
      if (e != hipSuccess) {
        S.uploaded = nullptr;
        return hipfail(ctx, e, "batch table event");
      }
    }
    if (S.cap >= need) continue;
    const size_t cap = std::max<size_t>(need, 2 * S.cap);
    HIPCHK(ctx, hipEventSynchronize(S.uploaded));  // the last upload from the old pinned copy
    HIPCHK(ctx, hipStreamSynchronize(stream_of(ctx)));  // kernels reading the old device table
    if (S.dtab) hipFree(S.dtab);
    if (S.htab) hipHostFree(S.htab);
    S.dtab = S.htab = nullptr;
    S.cap = 0;
    S.valid = false;
    hipError_t e = hipMalloc(&S.dtab, cap * sizeof(BatchEntry));
    if (e == hipSuccess) e = hipHostMalloc((void**)&S.htab, cap * sizeof(BatchEntry), hipHostMallocDefault);
    if (e != hipSuccess) return hipfail(ctx, e, "batch table allocation");
    S.cap = cap;
  }
  b->dgs.assign(members, members + n);
  b->groups.swap(groups);
  return ATM_OK;
}

}  // namespace atm

extern "C" {

int atm_batch_create(atm_ctx* ctx, int32_t n, atm_dg* const* members, atm_batch** out) {
  if (!ctx || !out || n < 0 || (n > 0 && !members)) return fail(ctx, ATM_ERR_ARG, "atm_batch_create: bad argument");
  *out = nullptr;
  atm_batch* b = new atm_batch();
  b->ctx = ctx;
  const int rc = batch_set_members(b, n, members);
  if (rc != ATM_OK) {
    atm_batch_destroy(b);
    return rc;
  }
  *out = b;
  return ATM_OK;
}

int atm_batch_destroy(atm_batch* b) {
  if (!b) return ATM_ERR_ARG;
  hipSetDevice(b->ctx->device);
  hipStreamSynchronize(stream_of(b->ctx));
  for (hipGraphExec_t& x : b->exec)
    if (x) hipGraphExecDestroy(x);
  for (auto& S : b->slots) {
    if (S.uploaded) hipEventSynchronize(S.uploaded);
    if (S.dtab) hipFree(S.dtab);
    if (S.htab) hipHostFree(S.htab);
    if (S.uploaded) hipEventDestroy(S.uploaded);
  }
  delete b;
  return ATM_OK;
}

int atm_batch_ngroups(const atm_batch* b) { return b ? (int)b->groups.size() : ATM_ERR_ARG; }

int atm_batch_lsrk_run(atm_batch* b, int scheme, double* const* q, double* const* dq, double* const* scratch,
                       const double* dt, int64_t nsteps, int use_graph) {
  if (!b || !q || !dq || !scratch || !dt) return fail(b ? b->ctx : nullptr, ATM_ERR_ARG, "atm_batch_lsrk_run: null argument");
  if (scheme != ATM_LSRK54 && scheme != ATM_LSRK144) return fail(b->ctx, ATM_ERR_ARG, "unknown LSRK scheme");
  atm_ctx* ctx = b->ctx;
  hipSetDevice(ctx->device);
  const int n = (int)b->dgs.size();
  for (int i = 0; i < n; ++i) {
    if (!q[i] || !dq[i] || !scratch[i] || q[i] == scratch[i] || q[i] == dq[i] || dq[i] == scratch[i])
      return fail(ctx, ATM_ERR_ARG, "atm_batch_lsrk_run: member buffers missing or aliased");
  }
  // device tables (flattened element table: first workgroup of each member).  The
  // table is uploaded from the group's pinned host copy, and only when it changed
  // (members' buffers, time steps or shapes); before the host copy is rewritten,
  // the previous upload from it must have finished.
  for (auto& G : b->groups) {
    atm_batch::Slot& S = b->slots[G.slot];
    std::vector<BatchEntry> tab(G.members.size());
    std::memset(tab.data(), 0, tab.size() * sizeof(BatchEntry));  // padding too: compared bytewise
    int64_t start = 0;
    for (size_t k = 0; k < G.members.size(); ++k) {
      const int i = G.members[k];
      const atm_dg* dg = b->dgs[i];
      tab[k].p = dg->dev;
      const int e = G.multi ? G.mepw[k] : G.epw;
      if (G.multi) {
        tab[k].p.epw = e;
        tab[k].p.lds = G.mlds[k];
      }
      tab[k].buf[0] = q[i];
      tab[k].buf[1] = scratch[i];
      tab[k].dq = dq[i];
      tab[k].dt = dt[i];
      tab[k].block_start = start;
      start += (dg->dev.K + e - 1) / e;
    }
    G.blocks = start;
    const size_t bytes = tab.size() * sizeof(BatchEntry);
    if (S.valid && S.used == tab.size() && std::memcmp(S.htab, tab.data(), bytes) == 0) continue;
    HIPCHK(ctx, hipEventSynchronize(S.uploaded));
    std::memcpy(S.htab, tab.data(), bytes);
    HIPCHK(ctx, hipMemcpyAsync(S.dtab, S.htab, bytes, hipMemcpyHostToDevice, stream_of(ctx)));
    HIPCHK(ctx, hipEventRecord(S.uploaded, stream_of(ctx)));
    S.valid = true;
    S.used = tab.size();
  }
  const LSRKTable t = lsrk_table(scheme);
  int parity = 0;
  auto one_step = [&]() -> int {
    for (int st = 0; st < t.ns; ++st) {
      for (auto& G : b->groups) {
        if (G.blocks == 0) continue;
        hipError_t e = G.bfn(b->slots[G.slot].dtab, (int)G.members.size(), G.blocks, G.threads, G.lds, G.epw,
                             stream_of(ctx), parity,
                             t.a[st], t.a[st] == 0.0 ? 1 : 0, t.b[st]);
        if (e != hipSuccess) return hipfail(ctx, e, "batched element kernel launch");
      }
      parity = 1 - parity;
    }
    return ATM_OK;
  };
  const int64_t pair = (t.ns % 2 == 1) ? 2 : 1;
  int64_t done = 0;
  if (use_graph && nsteps >= pair) {
    // the captured kernels read the member tables, so one instantiated graph per
    // scheme serves every later call on this batch
    hipStream_t s = stream_of(ctx);
    hipGraphExec_t& exec = b->exec[scheme == ATM_LSRK54 ? 0 : 1];
    if (!exec) {
      hipGraph_t graph = nullptr;
      HIPCHK(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      int rc = ATM_OK;
      for (int64_t i = 0; i < pair && rc == ATM_OK; ++i) rc = one_step();
      hipError_t ec = hipStreamEndCapture(s, &graph);
      if (rc) {
        if (graph) hipGraphDestroy(graph);
        return rc;
      }
      if (ec != hipSuccess) return hipfail(ctx, ec, "hipStreamEndCapture");
      hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      hipGraphDestroy(graph);
      if (ei != hipSuccess) {
        exec = nullptr;
        return hipfail(ctx, ei, "hipGraphInstantiate");
      }
    }
    const int64_t reps = nsteps / pair;
    for (int64_t i = 0; i < reps; ++i) HIPCHK(ctx, hipGraphLaunch(exec, s));
    done = reps * pair;
  }
  for (; done < nsteps; ++done) {
    int rc = one_step();
    if (rc) return rc;
  }
  if (parity != 0) {
    for (int i = 0; i < n; ++i) {
      const int64_t m = (int64_t)b->dgs[i]->Ns * b->dgs[i]->Ntot;
      hipLaunchKernelGGL(copy_kernel, dim3(ew_blocks(m)), dim3(EW_THREADS), 0, stream_of(ctx), q[i], scratch[i], m);
    }
    HIPCHK(ctx, hipGetLastError());
  }
  return ATM_OK;
}

// ---- banded -----------------------------------------------------------------
int atm_banded_create(atm_dg* lin, int32_t KV, const int32_t* colelem, int32_t ncols, double scale,
                      atm_banded** out) {
  if (!lin || !colelem || !out || KV < 1 || ncols < 0)
    return fail(lin ? lin->ctx : nullptr, ATM_ERR_ARG, "atm_banded_create: bad argument");
  atm_ctx* ctx = lin->ctx;
  *out = nullptr;
  if (lin->dev.dirmask != (1 << (lin->dim - 1)))
    return fail(ctx, ATM_ERR_ARG, "banded operator must act in the last direction only");
  const int64_t K = lin->dev.K;
  if ((int64_t)KV * ncols != K) return fail(ctx, ATM_ERR_ARG, "columns must cover every element once");
  std::vector<int32_t> col(K, -1), ez(K, 0);
  for (int64_t c = 0; c < ncols; ++c)
    for (int z = 0; z < KV; ++z) {
      const int32_t e = colelem[c * KV + z];
      if (e < 0 || e >= K || col[e] != -1) return fail(ctx, ATM_ERR_ARG, "colelem is not a permutation");
      col[e] = (int32_t)c;
      ez[e] = z;
    }
  hipSetDevice(ctx->device);
  atm_banded* b = new atm_banded();
  b->lin = lin;
  b->KV = KV;
  b->ncols = ncols;
  b->Nh = lin->Np / lin->Nq;
  b->n = lin->Ns * lin->Nq * KV;
  b->bw = lin->Ns * lin->Nq - 1;
  b->nb = 2 * b->bw + 1;
  b->nlines = (int64_t)ncols * b->Nh;
  auto cleanup = [&](int rc) {
    atm_banded_destroy(b);
    return rc;
  };
  const size_t bandsz = (size_t)b->nlines * b->n * b->nb;
  hipError_t e;
  double* band = nullptr;  // assembled band of every line, [(r*nb + j) * nlines + line]
  if ((e = hipMalloc(&band, bandsz * sizeof(double) + 8)) != hipSuccess) return cleanup(hipfail(ctx, e, "hipMalloc band"));
  if ((e = hipMalloc(&b->elem_col, K * sizeof(int32_t) + 8)) != hipSuccess) return cleanup(hipfail(ctx, e, "hipMalloc"));
  if ((e = hipMalloc(&b->elem_ez, K * sizeof(int32_t) + 8)) != hipSuccess) return cleanup(hipfail(ctx, e, "hipMalloc"));
  if ((e = hipMalloc(&b->colelem, K * sizeof(int32_t) + 8)) != hipSuccess) return cleanup(hipfail(ctx, e, "hipMalloc"));
  b->dev_bytes += 3 * (K * sizeof(int32_t) + 8);
  if (K > 0) {
    hipMemcpy(b->elem_col, col.data(), K * sizeof(int32_t), hipMemcpyHostToDevice);
    hipMemcpy(b->elem_ez, ez.data(), K * sizeof(int32_t), hipMemcpyHostToDevice);
    hipMemcpy(b->colelem, colelem, K * sizeof(int32_t), hipMemcpyHostToDevice);
  }
  const int64_t nfield = (int64_t)lin->Ns * lin->Ntot;
  double *pq = nullptr, *pdq = nullptr;
  if ((e = hipMalloc(&pq, nfield * sizeof(double) + 8)) != hipSuccess) {
    hipFree(band);
    return cleanup(hipfail(ctx, e, "hipMalloc probe"));
  }
  if ((e = hipMalloc(&pdq, nfield * sizeof(double) + 8)) != hipSuccess) {
    hipFree(pq);
    hipFree(band);
    return cleanup(hipfail(ctx, e, "hipMalloc probe"));
  }
  const BandGeo g = band_geo(b);
  hipStream_t s = stream_of(ctx);
  hipLaunchKernelGGL(fill_kernel, dim3(ew_blocks(bandsz)), dim3(EW_THREADS), 0, s, band, 0.0, (int64_t)bandsz);
  int rc = ATM_OK;
  for (int color = 0; color < b->nb && rc == ATM_OK; ++color) {
    hipLaunchKernelGGL(band_probe_kernel, dim3(ew_blocks(nfield)), dim3(EW_THREADS), 0, s, g, pq, color);
    rc = launch(lin, pq, pdq, 1.0, 1, nullptr, 0.0);
    if (rc == ATM_OK)
      hipLaunchKernelGGL(band_scatter_kernel, dim3(ew_blocks(b->nlines * b->n)), dim3(EW_THREADS), 0, s, g,
                         pdq, band, color);
  }
  hipFree(pq);
  hipFree(pdq);
  if (rc) {
    hipFree(band);
    return cleanup(rc);
  }
  // A = I + scale*L, then keep one copy of each distinct line matrix
  hipLaunchKernelGGL(band_scale_kernel, dim3(ew_blocks(bandsz)), dim3(EW_THREADS), 0, s, g, band, scale);
  const int64_t L = b->nlines;
  uint64_t* dhash = nullptr;
  int64_t* drep = nullptr;
  int32_t* dsame = nullptr;
  std::vector<uint64_t> hash(L);
  std::vector<int64_t> rep(L);
  std::vector<int32_t> same(L);
  auto done = [&](int code) {
    if (dhash) hipFree(dhash);
    if (drep) hipFree(drep);
    if (dsame) hipFree(dsame);
    hipFree(band);
    return code ? cleanup(code) : ATM_OK;
  };
  if ((e = hipMalloc(&dhash, L * sizeof(uint64_t) + 8)) != hipSuccess) return done(hipfail(ctx, e, "hipMalloc"));
  if ((e = hipMalloc(&drep, L * sizeof(int64_t) + 8)) != hipSuccess) return done(hipfail(ctx, e, "hipMalloc"));
  if ((e = hipMalloc(&dsame, L * sizeof(int32_t) + 8)) != hipSuccess) return done(hipfail(ctx, e, "hipMalloc"));
  const int64_t lblocks = (L + 63) / 64;
  const int nchunk = (int)(((int64_t)b->n * b->nb + HASH_CHUNK - 1) / HASH_CHUNK);
  uint64_t* dpart = nullptr;
  if ((e = hipMalloc(&dpart, (size_t)nchunk * L * sizeof(uint64_t) + 8)) != hipSuccess)
    return done(hipfail(ctx, e, "hipMalloc"));
  if (lblocks > 0) {
    hipLaunchKernelGGL(band_hash_chunk_kernel, dim3(lblocks, nchunk), dim3(64), 0, s, g, band, dpart);
    hipLaunchKernelGGL(band_hash_combine_kernel, dim3(lblocks), dim3(64), 0, s, g, dpart, nchunk, dhash);
  }
  if ((e = hipMemcpyAsync(hash.data(), dhash, L * sizeof(uint64_t), hipMemcpyDeviceToHost, s)) != hipSuccess ||
      (e = hipStreamSynchronize(s)) != hipSuccess) {
    hipFree(dpart);
    return done(hipfail(ctx, e, "banded assembly"));
  }
  hipFree(dpart);
  {
    std::vector<std::pair<uint64_t, int64_t>> first;
    first.reserve(64);
    for (int64_t l = 0; l < L; ++l) {
      int64_t r0 = -1;
      for (auto& pr : first)
        if (pr.first == hash[l]) { r0 = pr.second; break; }
      if (r0 < 0) {
        if (first.size() < 4096) first.push_back({hash[l], l});
        r0 = l;
      }
      rep[l] = r0;
    }
  }
  if ((e = hipMemcpyAsync(drep, rep.data(), L * sizeof(int64_t), hipMemcpyHostToDevice, s)) != hipSuccess)
    return done(hipfail(ctx, e, "hipMemcpy"));
  if (lblocks > 0) {
    hipLaunchKernelGGL(fill_i32_kernel, dim3(lblocks), dim3(64), 0, s, dsame, 1, L);
    hipLaunchKernelGGL(band_verify_chunk_kernel, dim3(lblocks, nchunk), dim3(64), 0, s, g, band, drep, dsame);
  }
  if ((e = hipMemcpyAsync(same.data(), dsame, L * sizeof(int32_t), hipMemcpyDeviceToHost, s)) != hipSuccess ||
      (e = hipStreamSynchronize(s)) != hipSuccess)
    return done(hipfail(ctx, e, "banded verification"));
  std::vector<int64_t> src;
  std::vector<int64_t> uid(L, -1);
  b->line_mat_host.assign(L, 0);
  for (int64_t l = 0; l < L; ++l) {
    const int64_t r0 = same[l] ? rep[l] : l;
    if (uid[r0] < 0) {
      uid[r0] = (int64_t)src.size();
      src.push_back(r0);
    }
    b->line_mat_host[l] = (int32_t)uid[r0];
  }
  b->nu = (int64_t)src.size();
  // lines grouped by matrix (ascending line order within a group), each group padded to whole
  // runs of 64 lines by repeating the run's first line (the repeated lines compute the same
  // values and store them to the same place)
  std::vector<int32_t> perm, wmat;
  {
    std::vector<std::vector<int32_t>> by(b->nu);
    for (int64_t l = 0; l < L; ++l) by[b->line_mat_host[l]].push_back((int32_t)l);
    for (int64_t u = 0; u < b->nu; ++u) {
      for (size_t i = 0; i < by[u].size(); i += 64) {
        for (size_t k = i; k < i + 64; ++k) perm.push_back(k < by[u].size() ? by[u][k] : by[u][i]);
        wmat.push_back((int32_t)u);
      }
    }
  }
  b->nwaves = (int64_t)wmat.size();
  // runs of 16 lines for the block-form solves; a run's padding lanes repeat its first line
  // (they compute that line's values again and store them to the same place)
  std::vector<int32_t> perm16, wmat16;
  {
    std::vector<std::vector<int32_t>> by(b->nu);
    for (int64_t l = 0; l < L; ++l) by[b->line_mat_host[l]].push_back((int32_t)l);
    for (int64_t u = 0; u < b->nu; ++u)
      for (size_t i = 0; i < by[u].size(); i += 16) {
        for (size_t k = i; k < i + 16; ++k) perm16.push_back(k < by[u].size() ? by[u][k] : by[u][i]);
        wmat16.push_back((int32_t)u);
      }
  }
  b->nwaves16 = (int64_t)wmat16.size();
  if ((e = hipMalloc(&b->perm16, perm16.size() * sizeof(int32_t) + 8)) != hipSuccess ||
      (e = hipMalloc(&b->wave_mat16, wmat16.size() * sizeof(int32_t) + 8)) != hipSuccess)
    return done(hipfail(ctx, e, "hipMalloc"));
  b->dev_bytes += (perm16.size() + wmat16.size()) * sizeof(int32_t) + 16;
  if (!perm16.empty() &&
      ((e = hipMemcpy(b->perm16, perm16.data(), perm16.size() * sizeof(int32_t), hipMemcpyHostToDevice)) != hipSuccess ||
       (e = hipMemcpy(b->wave_mat16, wmat16.data(), wmat16.size() * sizeof(int32_t), hipMemcpyHostToDevice)) !=
           hipSuccess))
    return done(hipfail(ctx, e, "hipMemcpy"));
  if ((e = hipMalloc(&b->perm, perm.size() * sizeof(int32_t) + 8)) != hipSuccess ||
      (e = hipMalloc(&b->wave_mat, wmat.size() * sizeof(int32_t) + 8)) != hipSuccess)
    return done(hipfail(ctx, e, "hipMalloc"));
  b->dev_bytes += (perm.size() + wmat.size()) * sizeof(int32_t) + 16;
  if (!perm.empty() && ((e = hipMemcpy(b->perm, perm.data(), perm.size() * sizeof(int32_t), hipMemcpyHostToDevice)) != hipSuccess ||
                        (e = hipMemcpy(b->wave_mat, wmat.data(), wmat.size() * sizeof(int32_t), hipMemcpyHostToDevice)) != hipSuccess))
    return done(hipfail(ctx, e, "hipMemcpy"));
  const size_t facsz = (size_t)b->n * b->nb * b->nu;
  if ((e = hipMalloc(&b->fac, facsz * sizeof(double) + 8)) != hipSuccess) return done(hipfail(ctx, e, "hipMalloc factors"));
  if ((e = hipMalloc(&b->line_mat, L * sizeof(int32_t) + 8)) != hipSuccess) return done(hipfail(ctx, e, "hipMalloc"));
  b->dev_bytes += facsz * sizeof(double) + 8 + L * sizeof(int32_t) + 8;
  if ((e = hipMemcpyAsync(drep, src.data(), src.size() * sizeof(int64_t), hipMemcpyHostToDevice, s)) != hipSuccess ||
      (e = hipMemcpyAsync(b->line_mat, b->line_mat_host.data(), L * sizeof(int32_t), hipMemcpyHostToDevice, s)) !=
          hipSuccess)
    return done(hipfail(ctx, e, "hipMemcpy"));
  if (facsz > 0)
    hipLaunchKernelGGL(band_compact_kernel, dim3(ew_blocks(facsz)), dim3(EW_THREADS), 0, s, g, band, drep, b->nu,
                       b->fac);
  if (b->nu > 0) {
    const size_t lu_lds = (size_t)b->n * b->nb * sizeof(double);
    if (lu_lds <= ctx->max_lds)
      hipLaunchKernelGGL(band_lu_wg_kernel<true>, dim3(b->nu), dim3(LU_THREADS), lu_lds, s, g, b->fac, b->nu);
    else
      hipLaunchKernelGGL(band_lu_wg_kernel<false>, dim3(b->nu), dim3(LU_THREADS), 0, s, g, b->fac, b->nu);
    if ((e = hipGetLastError()) != hipSuccess) return done(hipfail(ctx, e, "band LU launch"));
    // block form for the block solve (up to 2 GiB; beyond that the register-blocked substitution)
    const size_t blk_bytes = (size_t)b->nu * b->KV * 4 * (size_t)(b->bw + 1) * (b->bw + 1) * sizeof(double);
    if (band_blk_kernel(lin->Ns, lin->Nq) && blk_bytes <= ((size_t)1 << 31)) {
      if ((e = hipMalloc(&b->blocks, blk_bytes)) != hipSuccess) return done(hipfail(ctx, e, "hipMalloc blocks"));
      b->dev_bytes += blk_bytes;
      hipLaunchKernelGGL(band_blocks_kernel, dim3(b->nu * b->KV), dim3(64), 0, s, g, b->fac, b->blocks);
      if ((e = hipGetLastError()) != hipSuccess) return done(hipfail(ctx, e, "band blocks launch"));
      if (band_mfma_kernel(lin->Ns, lin->Nq, 1)) {
        const int BS = b->bw + 1, KT = (BS + 3) / 4;
        const size_t fbytes = ((size_t)b->nu * b->KV * 4 * KT * KT * 16 + band_mfma_frag_pad()) * sizeof(double);
        if ((e = hipMalloc(&b->frags, fbytes)) != hipSuccess) return done(hipfail(ctx, e, "hipMalloc fragments"));
        b->dev_bytes += fbytes;
        hipLaunchKernelGGL(band_frags_kernel, dim3(b->nu * b->KV), dim3(256), 0, s, g, b->blocks, b->frags, KT);
        if ((e = hipGetLastError()) != hipSuccess) return done(hipfail(ctx, e, "band fragments launch"));
      }
    }
  }
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return done(hipfail(ctx, e, "banded factorisation"));
  rc = done(ATM_OK);
  if (rc) return rc;
  *out = b;
  return ATM_OK;
}

int64_t atm_banded_device_bytes(const atm_banded* b) { return b ? (int64_t)b->dev_bytes : ATM_ERR_ARG; }

int atm_banded_destroy(atm_banded* b) {
  if (!b) return ATM_ERR_ARG;
  if (b->fac) hipFree(b->fac);
  if (b->line_mat) hipFree(b->line_mat);
  if (b->perm) hipFree(b->perm);
  if (b->wave_mat) hipFree(b->wave_mat);
  if (b->blocks) hipFree(b->blocks);
  if (b->frags) hipFree(b->frags);
  if (b->perm16) hipFree(b->perm16);
  if (b->wave_mat16) hipFree(b->wave_mat16);
  if (b->elem_col) hipFree(b->elem_col);
  if (b->elem_ez) hipFree(b->elem_ez);
  if (b->colelem) hipFree(b->colelem);
  delete b;
  return ATM_OK;
}

}  // extern "C"

namespace {

// Waves per workgroup of the matrix-core solve for this factorisation: 4 (64 lines of a
// matrix sharing each block's fragments in LDS) when that LDS fits, else 1; 0 when the
// (Ns, Nq) is not instantiated or not even one wave's LDS fits.
bool band_mfma_fits(const atm_banded* b, int W) {
  return band_mfma_kernel(b->lin->Ns, b->lin->Nq, W) &&
         band_mfma_lds_bytes(b->n, b->KV, b->bw + 1, W) <= b->lin->ctx->max_lds;
}

int band_mfma_waves(const atm_banded* b) {
  if (b->mfma_w > 0 && band_mfma_fits(b, b->mfma_w)) return b->mfma_w;  // chosen (set or tuned)
  for (int W : {4, 1})
    if (band_mfma_fits(b, W)) return W;
  return 0;
}

// Whether solve kernel k can run this factorisation (operand tables built, the
// (Ns, Nq) instantiated, LDS large enough).
bool band_kernel_ok(const atm_banded* b, int k) {
  const atm_ctx* ctx = b->lin->ctx;
  switch (k) {
    case ATM_BAND_MFMA:
      return b->frags && band_mfma_waves(b) > 0;
    case ATM_BAND_BLOCK:
      return b->blocks && band_blk_kernel(b->lin->Ns, b->lin->Nq);
    case ATM_BAND_REG:
      return band_reg_kernel(b->lin->Ns, b->lin->Nq) != nullptr;
    case ATM_BAND_LDS:
      return band_solve_lds_bytes(b->n) <= ctx->max_lds;
    case ATM_BAND_PLAIN:
      return true;
    default:
      return false;
  }
}

// The kernel the automatic choice takes: the first available in the order of
// measured speed on the configurations here (MFMA block form, register block
// form, register-blocked substitution, LDS substitution, plain substitution).
int band_auto_kernel(const atm_banded* b) {
  for (int k : {ATM_BAND_MFMA, ATM_BAND_BLOCK, ATM_BAND_REG, ATM_BAND_LDS})
    if (band_kernel_ok(b, k)) return k;
  return ATM_BAND_PLAIN;
}

// Solve kernel k on right-hand side rhs, or, with ark (kernel k the matrix-core solve and a
// fused instance for nk terms present), on the ARK stage combination formed in its row loads.
int band_launch(atm_banded* b, int k, double* x, const double* rhs, const ArkRhs* ark = nullptr, int nk = 0) {
  atm_ctx* ctx = b->lin->ctx;
  const BandGeo g = band_geo(b);
  hipStream_t st = stream_of(ctx);
  switch (k) {
    case ATM_BAND_MFMA: {
      const int W = band_mfma_waves(b);
      const size_t lds = band_mfma_lds_bytes(b->n, b->KV, b->bw + 1, W);
      if (W == 4 && b->nwaves > 0)
        hipLaunchKernelGGL(band_mfma_kernel(b->lin->Ns, b->lin->Nq, 4, nk), dim3(b->nwaves), dim3(256), lds, st, g,
                           b->frags, b->perm, b->wave_mat, x, rhs, ark ? *ark : ArkRhs{});
      else if (W == 1 && b->nwaves16 > 0)
        hipLaunchKernelGGL(band_mfma_kernel(b->lin->Ns, b->lin->Nq, 1, nk), dim3(b->nwaves16), dim3(64), lds, st, g,
                           b->frags, b->perm16, b->wave_mat16, x, rhs, ark ? *ark : ArkRhs{});
      break;
    }
    case ATM_BAND_BLOCK:
      if (b->nwaves16 > 0)
        hipLaunchKernelGGL(band_blk_kernel(b->lin->Ns, b->lin->Nq), dim3(b->nwaves16), dim3(64), 0, st, g, b->blocks,
                           b->perm16, b->wave_mat16, x, rhs);
      break;
    case ATM_BAND_REG:
      if (b->nwaves > 0)
        hipLaunchKernelGGL(band_reg_kernel(b->lin->Ns, b->lin->Nq), dim3(b->nwaves), dim3(64), 0, st, g, b->fac,
                           b->perm, b->wave_mat, x, rhs);
      break;
    case ATM_BAND_LDS: {
      constexpr int LINES = 64 / BAND_SOLVE_PARTS;
      const int64_t blocks = (b->nlines + LINES - 1) / LINES;
      if (blocks > 0)
        hipLaunchKernelGGL(band_solve_lds_kernel<BAND_SOLVE_PARTS>, dim3(blocks), dim3(64), band_solve_lds_bytes(b->n),
                           st, g, b->fac, b->line_mat, b->nu, x, rhs);
      break;
    }
    default: {
      const int64_t blocks = (b->nlines + 63) / 64;
      if (blocks > 0)
        hipLaunchKernelGGL(band_solve_kernel, dim3(blocks), dim3(64), 0, st, g, b->fac, b->line_mat, b->nu, x, rhs);
      break;
    }
  }
  HIPCHK(ctx, hipGetLastError());
  return ATM_OK;
}

}  // namespace

extern "C" {

int atm_banded_solve(atm_banded* b, double* x, const double* rhs) {
  if (!b || !x || !rhs) return ATM_ERR_ARG;
  return band_launch(b, b->kernel == ATM_BAND_AUTO ? band_auto_kernel(b) : b->kernel, x, rhs);
}

int atm_banded_get_kernel(const atm_banded* b) {
  return b ? (b->kernel == ATM_BAND_AUTO ? band_auto_kernel(b) : b->kernel) : ATM_ERR_ARG;
}

int atm_banded_set_kernel(atm_banded* b, int kernel) {
  if (!b) return ATM_ERR_ARG;
  if (kernel == ATM_BAND_AUTO) {
    b->kernel = kernel;
    return ATM_OK;
  }
  if (kernel < ATM_BAND_MFMA || kernel > ATM_BAND_PLAIN)
    return fail(b->lin->ctx, ATM_ERR_ARG, "atm_banded_set_kernel: unknown solve kernel");
  if (!band_kernel_ok(b, kernel))
    return fail(b->lin->ctx, ATM_ERR_UNSUPPORTED, "atm_banded_set_kernel: this kernel cannot solve this factorisation");
  b->kernel = kernel;
  return ATM_OK;
}

int atm_banded_set_waves(atm_banded* b, int waves) {
  if (!b) return ATM_ERR_ARG;
  if (waves != 0 && waves != 1 && waves != 4)
    return fail(b->lin->ctx, ATM_ERR_ARG, "atm_banded_set_waves: waves must be 0, 1 or 4");
  if (waves != 0 && !band_mfma_fits(b, waves))
    return fail(b->lin->ctx, ATM_ERR_UNSUPPORTED, "atm_banded_set_waves: that workgroup does not fit this factorisation");
  b->mfma_w = waves;
  return ATM_OK;
}

int atm_banded_get_waves(const atm_banded* b) { return b ? band_mfma_waves(b) : ATM_ERR_ARG; }

int atm_banded_autotune(atm_banded* b, double* x, const double* rhs, int reps, int32_t* kernel_out, float* us_out,
                        int max) {
  if (!b || !x || !rhs || reps < 1) return ATM_ERR_ARG;
  atm_ctx* ctx = b->lin->ctx;
  hipSetDevice(ctx->device);
  hipStream_t st = stream_of(ctx);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(ctx, hipEventCreate(&e0));
  hipError_t ee = hipEventCreate(&e1);
  if (ee != hipSuccess) {
    hipEventDestroy(e0);
    return hipfail(ctx, ee, "hipEventCreate");
  }
  int best = -1, n = 0, rc = ATM_OK;
  float best_us = 0.f;
  auto timed = [&](int k) -> float {
    rc = band_launch(b, k, x, rhs);  // warm-up
    if (rc) return 0.f;
    hipEventRecord(e0, st);
    for (int i = 0; i < reps && rc == ATM_OK; ++i) rc = band_launch(b, k, x, rhs);
    hipEventRecord(e1, st);
    hipEventSynchronize(e1);
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    return ms * 1e3f / reps;
  };
  for (int k = ATM_BAND_MFMA; k <= ATM_BAND_PLAIN && rc == ATM_OK; ++k) {
    if (!band_kernel_ok(b, k)) continue;
    float us = 0.f;
    if (k == ATM_BAND_MFMA) {  // its second shape dimension: waves per workgroup
      int best_w = 0;
      for (int W : {4, 1}) {
        if (!band_mfma_fits(b, W) || rc) continue;
        b->mfma_w = W;
        const float t = timed(k);
        if (best_w == 0 || t < us) {
          best_w = W;
          us = t;
        }
      }
      b->mfma_w = best_w;
    } else {
      us = timed(k);
    }
    if (rc) break;
    if (kernel_out && n < max) kernel_out[n] = k;
    if (us_out && n < max) us_out[n] = us;
    ++n;
    if (best < 0 || us < best_us) {
      best = k;
      best_us = us;
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc) return rc;
  b->kernel = best;
  return best;
}

int64_t atm_banded_nunique(const atm_banded* b) { return b ? b->nu : ATM_ERR_ARG; }

int atm_banded_shape(const atm_banded* b, int64_t* nlines, int32_t* n, int32_t* bw) {
  if (!b) return ATM_ERR_ARG;
  if (nlines) *nlines = b->nlines;
  if (n) *n = b->n;
  if (bw) *bw = b->bw;
  return ATM_OK;
}

// host layout [line][row][bw+col-row]
int atm_banded_download(atm_banded* b, double* host) {
  if (!b || !host) return ATM_ERR_ARG;
  atm_ctx* ctx = b->lin->ctx;
  const size_t sz = (size_t)b->nu * b->n * b->nb;
  std::vector<double> tmp(sz);
  HIPCHK(ctx, hipStreamSynchronize(stream_of(ctx)));
  HIPCHK(ctx, hipMemcpy(tmp.data(), b->fac, sz * sizeof(double), hipMemcpyDeviceToHost));
  for (int64_t line = 0; line < b->nlines; ++line) {
    const int64_t u = b->line_mat_host[line];
    for (int r = 0; r < b->n; ++r)
      for (int k = 0; k < b->nb; ++k)
        host[((size_t)line * b->n + r) * b->nb + k] = tmp[((size_t)u * b->n + r) * b->nb + k];
  }
  return ATM_OK;
}

// ---- ARK23 ------------------------------------------------------------------
int atm_ark23_run(atm_dg* ex, atm_dg* lin, atm_banded* band, double* q, double* work, double dt, double t0,
                  int64_t nsteps, int split_rhs, int paperversion) {
  (void)t0;
  if (!lin || !band || !q || !work) return fail(lin ? lin->ctx : nullptr, ATM_ERR_ARG, "atm_ark23_run: null argument");
  atm_ctx* ctx = lin->ctx;
  if (ex && (ex->Ntot != lin->Ntot || ex->Ns != lin->Ns)) return fail(ctx, ATM_ERR_ARG, "operators differ in shape");
  if (band->lin != lin) return fail(ctx, ATM_ERR_ARG, "band was not built from lin");
  hipSetDevice(ctx->device);
  const int64_t n = (int64_t)lin->Ns * lin->Ntot;
  // src/odesolvers.jl:284-304
  const double s2 = std::sqrt(2.0);
  const double a32 = paperversion ? (3 + 2 * s2) / 6 : 0.5;
  const double exa[3][3] = {{0, 0, 0}, {2 - s2, 0, 0}, {1 - a32, a32, 0}};
  const double exb[3] = {1 / (2 * s2), 1 / (2 * s2), 1 - 1 / s2};
  const double ima[3][3] = {{0, 0, 0}, {1 - 1 / s2, 1 - 1 / s2, 0}, {1 / (2 * s2), 1 / (2 * s2), 1 - 1 / s2}};
  double* Qhat = work;
  double* Qs[3] = {q, work + n, work + 2 * n};
  double* exK[3] = {work + 3 * n, work + 4 * n, work + 5 * n};
  double* imK[3] = {work + 6 * n, work + 7 * n, work + 8 * n};
  hipStream_t s = stream_of(ctx);
  const int blocks = ew_blocks(n);
  for (int64_t step = 0; step < nsteps; ++step) {
    for (int i = 0; i < 3; ++i) {
      if (i > 0) {
        ArkCoef c{};
        for (int k = 0; k < 3; ++k) {
          c.ea[k] = exa[i][k];
          c.ia[k] = ima[i][k];
        }
        const int kern = band->kernel == ATM_BAND_AUTO ? band_auto_kernel(band) : band->kernel;
        int rc;
        if (kern == ATM_BAND_MFMA && band_mfma_kernel(lin->Ns, lin->Nq, band_mfma_waves(band), i)) {
          // the first implicit stage's combination formed in the solve's row loads (no Qhat pass)
          const ArkRhs ark{q, {exK[0], exK[1], exK[2]}, {imK[0], imK[1], imK[2]}, c, dt, split_rhs ? 0 : 1};
          rc = band_launch(band, kern, Qs[i], nullptr, &ark, i);
        } else {
          hipLaunchKernelGGL(ark_combine_kernel, dim3(blocks), dim3(EW_THREADS), 0, s, Qhat, q, dt, i, c, exK[0],
                             exK[1], exK[2], imK[0], imK[1], imK[2], n, split_rhs ? 0 : 1);
          rc = atm_banded_solve(band, Qs[i], Qhat);
        }
        if (rc) return rc;
      }
      int rc;
      if (ex) {
        if ((rc = launch(ex, Qs[i], exK[i], 1.0, 1, nullptr, 0.0))) return rc;
      } else {
        hipLaunchKernelGGL(fill_kernel, dim3(blocks), dim3(EW_THREADS), 0, s, exK[i], 0.0, n);
      }
      if ((rc = launch(lin, Qs[i], imK[i], 1.0, 1, nullptr, 0.0))) return rc;
    }
    // q += dt*(b_i exK_i + b~_i imK_i), i = 1..3 (src/odesolvers.jl:382-384)
    ArkCoef c{};
    for (int k = 0; k < 3; ++k) {
      c.ea[k] = exb[k];
      c.ia[k] = exb[k];
    }
    hipLaunchKernelGGL(ark_combine_kernel, dim3(blocks), dim3(EW_THREADS), 0, s, q, q, dt, 3, c, exK[0], exK[1],
                       exK[2], imK[0], imK[1], imK[2], n, split_rhs ? 0 : 1);
    HIPCHK(ctx, hipGetLastError());
  }
  return ATM_OK;
}

}  // extern "C"
