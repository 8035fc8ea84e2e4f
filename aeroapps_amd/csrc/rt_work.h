// Device work layout of the discrete-ordinate radiance pipeline.
//
// A call processes nobs*nch items (one item = one column at one channel),
// item g = obs*nch + ch, in chunks of up to `cap` items.  Per-item arrays are
// indexed by the chunk-local item `it`; per-layer arrays are item-major
// ([it][layer][field]) so that the wavefront solving one item reads its layer
// data contiguously.
#pragma once
#include "aa_common.h"

namespace aa {

// Per-(item, layer) transmittance/beam record written by the setup stage.
enum TrField {
  TR_DTAU = 0,     // delta-M scaled layer optical thickness
  TR_TDUSER,       // exp(-dtau / mu_user)
  TR_AVSEC,        // average secant of the solar beam in the layer
  TR_ITRANS,       // solar beam transmittance to the layer top
  TR_TDMUBAR,      // exp(-dtau * avsec)
  TR_ITRANSU,      // itrans / mu_user
  TR_EMULTUP,      // upwelling beam multiplier at the user angle
  TR_SIGP,         // avsec + 1/mu_user
  TR_NFIELDS
};

// The active items of a Fourier term are kept in LBINS lists by layer count
// (list k at list + k * cap, count counts[k]), so that the boundary-value
// solve, which sweeps a wavefront's items over the largest of their layer
// counts, gets wavefronts of items of similar height.
constexpr int LBINS = 8;
__host__ __device__ inline int lbin(int L, int km) {
  const int b = (L - 1) * LBINS / km;
  return b < 0 ? 0 : (b >= LBINS ? LBINS - 1 : b);
}

struct RTWork {
  // ---- call description
  int km = 0;          // layers allocated per column (Fortran km)
  int nch = 0, nobs = 0;
  int nmom = 0;        // Legendre moments given per layer (nMom)
  int npol = 0;
  int N = 0, S = 1;    // streams per hemisphere, Stokes components
  int nmoments = 0;    // min(2N-1, nmom): Fourier terms and scaled moments
  int plane_parallel = 0;
  double missing = 1e15;
  int item0 = 0, cap = 0;   // first item of the chunk, chunk capacity
  const int* item_list = nullptr;   // optional: chunk item `it` is global item item_list[item0 + it]
  // ---- inputs (device, Fortran layouts of the driver arguments)
  const double* rot = nullptr;     // (km,nobs,nch)
  const double* depol = nullptr;   // (nch)
  const double* alpha = nullptr;   // (km,nobs,nch) or null
  const void* tau = nullptr;       // (km,nch,nobs)  double or float
  const void* ssa = nullptr;       // (km,nch,nobs)
  const void* pmom = nullptr;      // (km,nch,nobs,nmom,npol)
  int optics_f32 = 0;              // tau/ssa/pmom are float (device Mie output)
  int opt_obs0 = 0;                // first observation held in tau/ssa/pmom
  int opt_nobs = 0;                // observations held in tau/ssa/pmom (their nobs extent)
  const double* he = nullptr;      // (km+1,nobs) metres
  const double* albedo = nullptr;  // (nobs,nch); null for the MODIS BRDF surface
  // MODIS land surface (BRDF: Lambertian + Ross-thick + Li-sparse kernels);
  // when kernel_wt is set the surface is the BRDF, else Lambertian.
  const double* kernel_wt = nullptr;  // (3,nch,nobs): fiso, fgeo, fvol
  const double* kparam = nullptr;     // (2,nch,nobs): Li-sparse h/b, b/r
  // Cox-Munk ocean surface (BRDF: one glint kernel); set instead of kernel_wt
  const double* u10m = nullptr;       // (nobs) m/s
  const double* v10m = nullptr;       // (nobs) m/s
  const double* mr = nullptr;         // (nch) refractive index of water
  // polarised land surfaces (three Stokes components): the BPDF-NDVI kernel
  // after a Lambertian kernel of factor albedo(obs, ch) or, with kernel_wt,
  // after the MODIS kernels
  const double* bpdf = nullptr;       // (3,nch,nobs): refractive index, NDVI, scaling factor C
  double* br_out = nullptr;           // (nobs,nch) exact direct-bounce BRDF
  double* brq_out = nullptr;          // (nobs,nch) its Mueller elements (1,2) and (1,3) (polarised surfaces)
  double* bru_out = nullptr;
  const double* brdf_tab = nullptr;   // azimuth / scaling quadrature tables (BrTab)
  double* brdf = nullptr;             // [it][brdf_stride] Fourier components of the BRDF
  int brdf_stride = 0;
  double* wsa_cache = nullptr;        // [4]: h/b, b/r and Li-sparse white-sky albedo of the chunk's first item
  const double* sza = nullptr;     // (nobs)
  const double* raa = nullptr;
  const double* vza = nullptr;
  const double* flux = nullptr;    // (nch,nobs) or null (1)
  const int* nlayers = nullptr;    // (nobs) or null (km)
  const int* latch = nullptr;      // first global item of the SZA==0 nadir latch
  // ---- per item (chunk-local)
  int* status = nullptr;           // 0 ok, else the element is MISSING
  int* L_it = nullptr;             // layers of the item's column
  int* beam_cutoff = nullptr;
  int* testconv = nullptr;
  int* fsaved = nullptr;
  int* nfour = nullptr;            // Fourier terms done
  // the active (item, layer) pairs of the current Fourier term, numbered
  // densely: item it owns slots loff[it] .. loff[it] + L - 1, lown[slot] = it;
  // lcnt[M] is the number of slots of term M (filled where the term's item
  // list is built, so columns of different layer counts leave no idle lanes)
  int* lown = nullptr;
  // launch shapes of the Fourier-term kernels (0 default, 1 alternative):
  // lanes per item of the band solver, pair groups of the SAB/DAB kernel
  int shape_bvp = 0, shape_sab = 0;
  // host side: an event recorded on the launch stream once the Fourier loop
  // has enqueued term mid_term (nullptr: none); the next optics block of a
  // call starts on another stream when it completes
  hipEvent_t mid_event = nullptr;
  int mid_term = 0;
  int* loff = nullptr;
  int* lcnt = nullptr;
  double* csza = nullptr;
  double* usec = nullptr;
  double* tsb = nullptr;           // transmittance of the whole slant solar path (0 if cut off)
  double* ss = nullptr;            // [it][3] first-order single scatter at TOA
  double* db = nullptr;            // [it][3] first-order direct bounce at TOA
  double* stokes = nullptr;        // [it][3] Fourier-summed TOA Stokes vector
  // ---- per (item, layer)
  double* deltau = nullptr;        // [it][km] unscaled
  double* fac1 = nullptr;          // [it][km] 1 - f*omega
  double* ext = nullptr;           // [it][km] scaled layer optical thickness per km of depth
  double* exact = nullptr;         // [it][km][3] FO phase function x TMS factor
  double* og = nullptr;            // [it][km][LM][S][S] omega * scaled Greek
  unsigned* lsc = nullptr;         // [it][km] bit m: layer scatters at Fourier m
  double* tr = nullptr;            // [it][km][TR_NFIELDS]
  double* lev = nullptr;           // [it][km+1][4]: tauslant, suntau, atten, FO mult(enhanced)
  double* lostrans = nullptr;      // [it][km+1]
  double* lbuf = nullptr;          // [it][km][lstride] homogeneous + particular solutions
  double* bbuf = nullptr;          // [it][bstride] surface terms of the bottom layer
  int lstride = 0, bstride = 0;
  // ---- quadrature and PI tables of the quadrature streams (device)
  const double* quad = nullptr;    // [qs N][qw N][xq: M][l][i][S][S]
  // ---- mixed precision: FP32 phase-matrix contraction (SAB/DAB) by MFMA
  int precision = 0;               // 0 FP64, 1 mixed
  const float* ytab = nullptr;     // [M][2n^2 rows][2N*S*S] contraction weights
  const double* ytab64 = nullptr;  // ytab in FP64 (FP64 path)
  float* sab32 = nullptr;          // [it][km][2n^2] SAB rows then DAB rows (FP32)
  double* sab64 = nullptr;         // the same in FP64 (FP64 path)
  // ---- outputs (device, (nobs,nch) Fortran)
  double* radiance = nullptr;
  double* reflectance = nullptr;
  double* qout = nullptr;
  double* uout = nullptr;
  int* nfourier = nullptr;
};

// Global item (obs*nch + ch) of chunk-local item `it`.
__device__ __forceinline__ int global_item(const RTWork& W, int it) {
  return W.item_list ? W.item_list[W.item0 + it] : W.item0 + it;
}

// Offsets inside one lbuf record (per item and layer) for n = N*S.  The
// eigenvectors X+/X- are stored for the upper-hemisphere streams i < N only:
// those of stream i + N are X-/X+ of stream i, negated for Stokes components
// o >= 2 (REG/vlidort_solutions.f90:680-700), which xrow() applies.
struct LOff {
  int n, S, N;
  __host__ __device__ constexpr LOff(int N_, int S_) : n(N_ * S_), S(S_), N(N_) {}
  __host__ __device__ constexpr int xp(int i, int o, int k) const { return (i * S + o) * n + k; }   // i < N
  __host__ __device__ constexpr int xn(int i, int o, int k) const { return n * n + (i * S + o) * n + k; }
  // X+ (plus = true) or X- of stream i in [0, 2N), Stokes o: offset of its n
  // values and their sign
  __host__ __device__ constexpr int xrow(bool plus, int i, int o, double& sgn) const {
    const bool lo = i >= N;
    sgn = (lo && o >= 2) ? -1.0 : 1.0;
    return (plus != lo) ? xp(lo ? i - N : i, o, 0) : xn(lo ? i - N : i, o, 0);
  }
  __host__ __device__ constexpr int tdel(int k) const { return 2 * n * n + k; }
  __host__ __device__ constexpr int wu(int i, int o) const { return 2 * n * n + n + i * S + o; }
  __host__ __device__ constexpr int wl(int i, int o) const { return 2 * n * n + 3 * n + i * S + o; }
  // source coefficients: src[o] = sum_q x(q) * coef(o, q) + csrc(o), q < 2n (lcon then mcon)
  __host__ __device__ constexpr int coef(int o, int q) const { return 2 * n * n + 5 * n + o * 2 * n + q; }
  __host__ __device__ constexpr int csrc(int o) const { return 2 * n * n + 5 * n + 2 * n * S + o; }
  __host__ __device__ constexpr int size() const { return (2 * n * n + 5 * n + 2 * n * S + S + 1) & ~1; }
};

// Offsets inside one bbuf record: surface reflection terms of the bottom
// layer for the current Fourier term, per quadrature row i < N (the
// Lambertian surface has the same value in every row and only at Fourier 0).
struct BOff {
  int N, n, S;
  __host__ __device__ constexpr BOff(int N_, int n_, int S_ = 1) : N(N_), n(n_), S(S_) {}
  // R2 terms of BOA row (i, o1) for unknown k, beam R2 of the row, the BOA
  // source coefficients of user Stokes component o (q < 2n) and its beam term
  __host__ __device__ constexpr int r2p(int i, int o1, int k) const { return (i * S + o1) * n + k; }
  __host__ __device__ constexpr int r2m(int i, int o1, int k) const { return N * S * n + (i * S + o1) * n + k; }
  __host__ __device__ constexpr int bcoef(int o, int q) const { return 2 * N * S * n + o * 2 * n + q; }
  __host__ __device__ constexpr int r2beam(int i, int o1) const { return 2 * N * S * n + 2 * n * S + i * S + o1; }
  __host__ __device__ constexpr int bc(int o) const { return 2 * N * S * n + 2 * n * S + N * S + o; }
  __host__ __device__ constexpr int size() const { return 2 * N * S * n + 2 * n * S + N * S + S; }
};

// Offsets inside one polarised BRDF record (per item, three Stokes
// components): the 3 x 3 Mueller blocks q = 3 o1 + o2 of the Fourier
// components M < 2N (BRDF_F incident j -> reflected i, USER_BRDF_F j ->
// user), the first column (o2 = 0) of BRDF_F_0 (solar -> i; the solar flux
// vector is (1, 0, 0)) and the exact direct-bounce BRDF of every block.
struct BPOff {
  int N;
  __host__ __device__ constexpr BPOff(int N_) : N(N_) {}
  __host__ __device__ constexpr int f(int M, int i, int j, int q) const { return ((M * N + i) * N + j) * 9 + q; }
  __host__ __device__ constexpr int f0(int M, int i, int o1) const { return 18 * N * N * N + (M * N + i) * 3 + o1; }
  __host__ __device__ constexpr int uf(int M, int j, int q) const {
    return 18 * N * N * N + 6 * N * N + (M * N + j) * 9 + q;
  }
  __host__ __device__ constexpr int db(int q) const { return 18 * N * N * N + 24 * N * N + q; }
  __host__ __device__ constexpr int size() const { return (18 * N * N * N + 24 * N * N + 9 + 1) & ~1; }
};

// Offsets inside one BRDF record (per item, scalar): Fourier components
// M < 2N of the MODIS BRDF in VLIDORT's arrays (BRDF_F incident j -> reflected
// i, BRDF_F_0 solar -> i, USER_BRDF_F j -> user, USER_BRDF_F_0 solar -> user)
// and the exact direct-bounce BRDF.
struct BROff {
  int N;
  __host__ __device__ constexpr BROff(int N_) : N(N_) {}
  __host__ __device__ constexpr int f(int M, int i, int j) const { return (M * N + i) * N + j; }
  __host__ __device__ constexpr int f0(int M, int i) const { return 2 * N * N * N + M * N + i; }
  __host__ __device__ constexpr int uf(int M, int j) const { return 2 * N * N * N + 2 * N * N + M * N + j; }
  __host__ __device__ constexpr int uf0(int M) const { return 2 * N * N * N + 4 * N * N + M; }
  __host__ __device__ constexpr int db() const { return 2 * N * N * N + 4 * N * N + 2 * N; }
  __host__ __device__ constexpr int size() const { return (2 * N * N * N + 4 * N * N + 2 * N + 2) & ~1; }
};

// Host-built quadrature tables of the BRDF supplement for N streams and NAZ
// half-range azimuth streams (BRDF_QUADRATURE_GAUSSIAN; 25 for the MODIS
// set-up, 50 for Cox-Munk), the 24-stream white-sky quadrature and the N
// quadrature streams with their sines.
constexpr int BR_NAZ = 25, BR_NAZ_CM = 50, BR_NSC = 24;
struct BrTab {
  int N, A;
  __host__ __device__ constexpr BrTab(int N_, int A_ = BR_NAZ) : N(N_), A(A_) {}
  __host__ __device__ constexpr int x(int k) const { return k; }
  __host__ __device__ constexpr int cx(int k) const { return A + k; }
  __host__ __device__ constexpr int sx(int k) const { return 2 * A + k; }
  __host__ __device__ constexpr int a(int k) const { return 3 * A + k; }
  __host__ __device__ constexpr int caz(int M, int k) const { return (4 + M) * A + k; }   // M < 2N
  __host__ __device__ constexpr int sq(int i) const { return (4 + 2 * N) * A + i; }
  __host__ __device__ constexpr int ssin(int i) const { return (4 + 2 * N) * A + BR_NSC + i; }
  __host__ __device__ constexpr int ssw(int i) const { return (4 + 2 * N) * A + 2 * BR_NSC + i; }
  __host__ __device__ constexpr int qs(int i) const { return (4 + 2 * N) * A + 3 * BR_NSC + i; }
  __host__ __device__ constexpr int qsin(int i) const { return (4 + 2 * N) * A + 3 * BR_NSC + N + i; }
  __host__ __device__ constexpr int wsa_ross() const { return (4 + 2 * N) * A + 3 * BR_NSC + 2 * N; }
  // sine factors a_k sin(M x_k) (the full-range table of the polarised
  // surfaces only; zero in the half-range tables)
  __host__ __device__ constexpr int saz(int M, int k) const { return (4 + 2 * N) * A + 3 * BR_NSC + 2 * N + 1 + M * A + k; }
  __host__ __device__ constexpr int size() const { return (4 + 4 * N) * A + 3 * BR_NSC + 2 * N + 1; }
};

// Rows/columns of the SAB/DAB contraction table for (N, S): rows r*n+c (SAB)
// and n*n + r*n + c (DAB), columns (l*S + o)*S + o3 over l < 2N.
struct YOff {
  int n, K;
  __host__ __device__ constexpr YOff(int N_, int S_) : n(N_ * S_), K(2 * N_ * S_ * S_) {}
  __host__ __device__ constexpr int rows() const { return 2 * n * n; }
  __host__ __device__ constexpr int at(int M, int row, int k) const { return (M * rows() + row) * K + k; }
  __host__ __device__ constexpr int size(int LM) const { return LM * rows() * K; }
};

// Quadrature table layout for (N, S): qs[N], qw[N], then for each Fourier term
// M < 2N the normalised PI matrices xq[l][i][s1][s2] (l < 2N, i < N).
struct QOff {
  int N, S;
  __host__ __device__ constexpr QOff(int N_, int S_) : N(N_), S(S_) {}
  __host__ __device__ constexpr int qs(int i) const { return i; }
  __host__ __device__ constexpr int qw(int i) const { return N + i; }
  __host__ __device__ constexpr int xq(int M, int l, int i, int s1, int s2) const {
    return 2 * N + ((((M * 2 * N + l) * N + i) * S + s1) * S + s2);
  }
  __host__ __device__ constexpr int size() const { return 2 * N + 2 * N * 2 * N * N * S * S; }
};


// Launch shapes of the Fourier-term kernels.  Lanes per item of the band
// solver: 8 for n = N*S <= 8 (8 items per wavefront), the smallest power of
// two holding 2n rows above; the alternative the tuner may pick: 4 lanes (16
// items per wavefront) for n <= 4, 16 lanes (4 items) for 5 <= n <= 8.
__host__ __device__ constexpr int bvp_group(int n) {
  return n <= 8 ? 8 : (2 * n <= 16 ? 16 : (2 * n <= 32 ? 32 : 64));
}
__host__ __device__ constexpr int bvp_group_alt(int n) { return n <= 4 ? 4 : (n <= 8 ? 16 : bvp_group(n)); }
__host__ __device__ constexpr int bvp_slots(int n, int G) { return (3 * n + G - 1) / G; }
// SAB/DAB kernel: 16-pair groups per wavefront (4 when the A tiles fit in
// registers, else 1; the alternative: 1).
__host__ __device__ constexpr int sab_tiles(int N, int S) {
  return ((2 * N * S * N * S + 15) / 16) * ((2 * N * S * S + 3) / 4);
}
__host__ __device__ constexpr int sab_groups(int N, int S) { return sab_tiles(N, S) <= 32 ? 4 : 1; }
__host__ __device__ constexpr int sab_groups_alt(int, int) { return 1; }

// Launch shapes (0 default, 1 alternative) and work-chunk cap of one
// (streams, Stokes, precision) class, from the library's tuning table.
struct TuneEntry { int shape_bvp = 0, shape_sab = 0; long chunk_items = 0; };
TuneEntry tuning_for(int N, int S, int precision);

// Appends the wavefront's lanes with `take` set to a Fourier term's item
// lists (list + b * cap for layer-count bin b, counts[b]) and numbers their
// (item, layer) slots from *lcount, with one atomic per wavefront and bin
// instead of one per item.  Every lane of the wavefront must call it.
// Returns the item's first slot (taking lanes) or -1.
__device__ inline int append_items(const RTWork& W, int it, int L, bool take, int* counts, int* list, int* lcount) {
  if (__ballot(take) == 0ull) return -1;
  const int lane = (int)(threadIdx.x & 63u);
  const unsigned long long below = (1ull << lane) - 1ull;
  const int mine = take ? L : 0;
  int v = mine;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int u = __shfl_up(v, d, 64);
    if (lane >= d) v += u;
  }
  const int tot = __shfl(v, 63, 64);
  int base = 0;
  if (lane == 0) base = atomicAdd(lcount, tot);
  base = __shfl(base, 0, 64);
  const int off = base + v - mine;
  const int b = lbin(L, W.km);
#pragma unroll
  for (int k = 0; k < LBINS; ++k) {
    const unsigned long long m = __ballot(take && b == k);
    if (m == 0ull) continue;
    const int leader = __ffsll((long long)m) - 1;
    int pos = 0;
    if (lane == leader) pos = atomicAdd(counts + k, __popcll(m));
    pos = __shfl(pos, leader, 64);
    if (take && b == k) list[(size_t)k * W.cap + pos + __popcll(m & below)] = it;
  }
  if (take) W.loff[it] = off;
  return take ? off : -1;
}

}  // namespace aa
