// Chem_Mie optics query and per-tracer optics summation on MI355X (kernel a).
//
// Reference: CB/Chem_MieMod.F90:789-965 (Chem_MieQueryByInt),
//            CB/Chem_MieTableMod.F90:612-627 (RH hash),
//            GEO/VLIDORT_MieMod.F90:124-336 (VLIDORT_getAOPvector)
// (CB = src/Shared/GMAO_Shared/Chem_Base, GEO = src/Components/rtms/geo_vlidort).
//
// The reference computes in IEEE float32 with a fixed operation order; the
// kernels below use explicitly rounded float32 intrinsics (no contraction into
// fused multiply-adds) in that same order, so outputs are bit-identical.
//
// Layout: one thread per (column, layer) element; the Fortran arrays
// qm(km,nq,nobs), rh(km,nobs), tau(km,nch,nobs), pmom(km,nch,nobs,nmom,npol)
// have the layer index fastest, so a wavefront reads and writes 64 consecutive
// layers of the same column: every access is coalesced.  The tables (< 1 MB)
// stay in L2.
#include "aa_common.h"
#include "prof.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace aa {

struct SnapF {
  int irh, irhp1;   // 0-based table nodes
  float arh;
};

// RH snap onto the 991-slot hash (CB/Chem_MieMod.F90:860-867).  The reference
// evaluates rhUse = max(rh,0.) and then overwrites it with min(rh,0.99);
// negative RH is caught by the isnap < 1 clamp.
__device__ __forceinline__ SnapF rh_snap(const MieDev& T, int tr, float rh) {
  float rhuse = fminf(rh, 0.99f);
  float x = __fmul_rn(__fadd_rn(rhuse, 0.001f), 1000.0f);
  int isnap = (int)x;   // Fortran INT(): truncation toward zero
  if (isnap < 1) isnap = 1;
  SnapF s;
  s.arh = T.rha[tr * NRH_MAP + isnap - 1];
  s.irh = T.rhi[tr * NRH_MAP + isnap - 1] - 1;
  int nrh = T.nrh_t[tr];
  s.irhp1 = (s.irh + 1 > nrh - 1) ? nrh - 1 : s.irh + 1;
  return s;
}

// table(irh)*(1.-arh) + table(irhp1)*arh in float32, no contraction
__device__ __forceinline__ float lin(float lo, float hi, float arh) {
  return __fadd_rn(__fmul_rn(lo, __fsub_rn(1.0f, arh)), __fmul_rn(hi, arh));
}

constexpr int MAXPOL = 6;
constexpr int MAXTR = 32;

// VLIDORT_getAOPvector for a range of tracers [0, ntr_eff) (ntr_eff < nq only
// when the reference would have returned early with rc 666/777, in which case
// the partial, un-normalised sums are what the reference leaves behind).
//
// One wavefront per (column, channel), KC layers at a time (KC * nq <= 128):
//  1. lane = (layer, tracer), in rounds of 64: RH snap and the tracer's tau_i,
//     ssa_i, g_i (independent per tracer, so all of them at once), into LDS;
//  2. lane = layer: the tracer-order sums tau, ssa*tau, g*ssa*tau and the
//     normalisation, exactly in the reference's order;
//  3. lane = moment: the phase-moment sums.  The two table rows of a (tracer,
//     channel, node) are contiguous in the moment index, so every table read
//     is a 256-byte coalesced load from L2; the KC layers give KC independent
//     accumulations per tracer.  Per output element the order is the
//     reference's (tracer order; (pm*ssa_i)*tau_i).
constexpr int MIE_KC_MAX = 8, MIE_CAP = 128;   // layers per chunk, (layer, tracer) slots
__global__ void __launch_bounds__(64) get_aop_kernel(MieDev T, int km, int nobs, int nch, int nq,
                                                     const int* __restrict__ tracer_idx,
                                                     const int* __restrict__ chan_idx,
                                                     const float* __restrict__ qm, const float* __restrict__ rh,
                                                     int nmom, int npol, int ntr_eff, int normalize,
                                                     float* __restrict__ tau_o, float* __restrict__ ssa_o,
                                                     float* __restrict__ g_o, float* __restrict__ pmom_o) {
  __shared__ float4 info[MIE_CAP];     // (node, arh, ssa_i, tau_i) of slot (layer, tracer)
  __shared__ int4 rows[MIE_CAP];       // (lower row, upper row, P22/P44 tied, tracer found) of slot (layer, tracer)
  __shared__ float gin[MIE_CAP];       // g_i of slot (layer, tracer)
  __shared__ float stl[MIE_KC_MAX];    // ssa*tau of the chunk's layers (normalisation)
  const int lane = threadIdx.x;
  const int i = blockIdx.x / nch, n = blockIdx.x - (blockIdx.x / nch) * nch;
  const int ich = chan_idx[n];
  const int ntr = ntr_eff > 0 ? ntr_eff : 1;
  const int KC = max(1, min(MIE_KC_MAX, MIE_CAP / ntr));
  const long ostride = (long)nobs * nch * km;   // stride of one moment
  const long obase = ((long)i * nch + n) * km;
  const long rowstride = (long)T.npol * T.nmom;
  const int npass = (npol == 6) ? 4 : npol;
  for (int k0 = 0; k0 < km; k0 += KC) {
    const int kc = min(KC, km - k0);
    // ---- 1. slot = (layer, tracer)
    for (int e = lane; e < kc * ntr; e += 64) {
      const int kk = e / ntr, iq = e - (e / ntr) * ntr;
      if (iq < ntr_eff) {
        const int k = k0 + kk;
        const int tr = tracer_idx[iq];
        if (tr >= 0) {
          SnapF s = rh_snap(T, tr, rh[(long)i * km + k]);
          const long tb = ((long)tr * T.nch + ich) * T.nrh;
          const float bext_in = lin(T.bext[tb + s.irh], T.bext[tb + s.irhp1], s.arh);
          const float bsca_in = lin(T.bsca[tb + s.irh], T.bsca[tb + s.irhp1], s.arh);
          const float g_in = lin(T.g[tb + s.irh], T.g[tb + s.irhp1], s.arh);
          const float q = qm[((long)i * nq + iq) * km + k];
          info[e] = make_float4(__int_as_float(s.irh), s.arh, __fdiv_rn(bsca_in, bext_in), __fmul_rn(bext_in, q));
          rows[e] = make_int4((int)tb + s.irh, (int)tb + s.irhp1, (npol == 6 && T.npol_t[tr] == 4) ? 1 : 0, 1);
          gin[e] = g_in;
        } else {
          rows[e] = make_int4(0, 0, 0, 0);
        }
      }
    }
    __syncthreads();
    // ---- 2. lane = layer: sums in tracer order, normalisation
    //      (GEO/VLIDORT_MieMod.F90:320-334; g and pmom use the normalised ssa)
    if (lane < kc) {
      float tau = 0.0f, ssa = 0.0f, g = 0.0f;
      for (int iq = 0; iq < ntr_eff; ++iq) {
        if (rows[lane * ntr + iq].w == 0) continue;
        const float4 inf = info[lane * ntr + iq];
        const float ssa_ = inf.z, tau_ = inf.w;
        tau = __fadd_rn(tau, tau_);
        ssa = __fadd_rn(ssa, __fmul_rn(ssa_, tau_));
        g = __fadd_rn(g, __fmul_rn(__fmul_rn(gin[lane * ntr + iq], ssa_), tau_));
      }
      float st = 0.0f;
      if (normalize) {
        if (tau > 0.0f) ssa = __fdiv_rn(ssa, tau);
        st = __fmul_rn(ssa, tau);
        if (st > 0.0f) g = __fdiv_rn(g, st);
      }
      const long o = obase + k0 + lane;
      tau_o[o] = tau; ssa_o[o] = ssa; g_o[o] = g;
      stl[lane] = st;
    }
    __syncthreads();
    // ---- 3. lane = moment.  Polarisations run in pairs (a, b) where the
    //      spherical extension ties b to a's running sum (P22 = P11, P44 = P33).
    if (pmom_o != nullptr) {
      for (int m0 = 0; m0 < nmom; m0 += 64) {
        const int m = m0 + lane;
        const bool mon = m < nmom;
        const int mr = mon ? m : 0;
        for (int ps = 0; ps < npass; ++ps) {
          int pa = ps, pb = -1;
          if (npol == 6) {
            pa = (ps == 0) ? 0 : (ps == 1) ? 2 : (ps == 2) ? 1 : 3;
            pb = (ps == 0) ? 4 : (ps == 1) ? 5 : -1;
          }
          float acc_a[MIE_KC_MAX], acc_b[MIE_KC_MAX];
#pragma unroll
          for (int kk = 0; kk < MIE_KC_MAX; ++kk) { acc_a[kk] = 0.0f; acc_b[kk] = 0.0f; }
          for (int iq = 0; iq < ntr_eff; ++iq) {
            // tracer found and tied polarisations: the same for every layer
            const int4 r0 = rows[iq];
            if (r0.w == 0) continue;
            const bool sph = r0.z != 0;
#pragma unroll
            for (int kk = 0; kk < MIE_KC_MAX; ++kk) {
              if (kk >= kc) break;
              const float4 inf = info[kk * ntr + iq];
              const int4 rr = rows[kk * ntr + iq];
              const float* plo = T.pmom + (long)rr.x * rowstride + mr;
              const float* phi = T.pmom + (long)rr.y * rowstride + mr;
              const float pma = lin(plo[pa * T.nmom], phi[pa * T.nmom], inf.y);
              acc_a[kk] = __fadd_rn(acc_a[kk], __fmul_rn(__fmul_rn(pma, inf.z), inf.w));
              if (pb >= 0) {
                if (sph) {
                  acc_b[kk] = acc_a[kk];
                } else {
                  const float pmb = lin(plo[pb * T.nmom], phi[pb * T.nmom], inf.y);
                  acc_b[kk] = __fadd_rn(acc_b[kk], __fmul_rn(__fmul_rn(pmb, inf.z), inf.w));
                }
              }
            }
          }
          if (mon) {
#pragma unroll
            for (int kk = 0; kk < MIE_KC_MAX; ++kk) {
              if (kk >= kc) break;
              const float st = stl[kk];
              if (normalize && st > 0.0f) { acc_a[kk] = __fdiv_rn(acc_a[kk], st); acc_b[kk] = __fdiv_rn(acc_b[kk], st); }
            }
            float* oa = pmom_o + ((long)pa * nmom + m) * ostride + obase + k0;
            float* ob = pb >= 0 ? pmom_o + ((long)pb * nmom + m) * ostride + obase + k0 : nullptr;
            // a moment's layers are contiguous: 16-byte stores when the chunk
            // holds whole aligned groups of 4 layers
            if (kc % 4 == 0 && (((size_t)oa | (size_t)ob) & 15) == 0) {
#pragma unroll
              for (int q = 0; q < MIE_KC_MAX / 4; ++q) {
                if (4 * q >= kc) break;
                *reinterpret_cast<float4*>(oa + 4 * q) =
                    make_float4(acc_a[4 * q], acc_a[4 * q + 1], acc_a[4 * q + 2], acc_a[4 * q + 3]);
                if (ob)
                  *reinterpret_cast<float4*>(ob + 4 * q) =
                      make_float4(acc_b[4 * q], acc_b[4 * q + 1], acc_b[4 * q + 2], acc_b[4 * q + 3]);
              }
            } else {
#pragma unroll
              for (int kk = 0; kk < MIE_KC_MAX; ++kk) {
                if (kk >= kc) break;
                oa[kk] = acc_a[kk];
                if (ob) ob[kk] = acc_b[kk];
              }
            }
          }
        }
      }
    }
    __syncthreads();
  }
}

// Chem_MieQuery over n elements of one tracer / table channel.
__global__ void mie_query_kernel(MieDev T, int tr, int ich, int n, const float* __restrict__ q,
                                 const float* __restrict__ rh, float* tau, float* ssa, float* gasym, float* bext,
                                 float* bsca, float* bbck, float* refr, float* refi, int nmom, int npol,
                                 float* pmom) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  SnapF s = rh_snap(T, tr, rh[e]);
  const long tb = ((long)tr * T.nch + ich) * T.nrh;
  float bext_in = lin(T.bext[tb + s.irh], T.bext[tb + s.irhp1], s.arh);
  float bsca_in = lin(T.bsca[tb + s.irh], T.bsca[tb + s.irhp1], s.arh);
  if (tau) tau[e] = __fmul_rn(bext_in, q[e]);
  if (ssa) ssa[e] = __fdiv_rn(bsca_in, bext_in);
  if (bext) bext[e] = bext_in;
  if (bsca) bsca[e] = bsca_in;
  if (gasym) gasym[e] = lin(T.g[tb + s.irh], T.g[tb + s.irhp1], s.arh);
  if (bbck) bbck[e] = lin(T.bbck[tb + s.irh], T.bbck[tb + s.irhp1], s.arh);
  if (refr) refr[e] = lin(T.refr[tb + s.irh], T.refr[tb + s.irhp1], s.arh);
  if (refi) refi[e] = lin(T.refi[tb + s.irh], T.refi[tb + s.irhp1], s.arh);
  if (pmom) {
    const float* plo = T.pmom + (tb + s.irh) * (long)T.npol * T.nmom;
    const float* phi = T.pmom + (tb + s.irhp1) * (long)T.npol * T.nmom;
    for (int p = 0; p < npol; ++p)
      for (int m = 0; m < nmom; ++m)
        pmom[((long)p * nmom + m) * n + e] = lin(plo[p * T.nmom + m], phi[p * T.nmom + m], s.arh);
  }
}

}  // namespace aa

// ====================================================================== host side
struct aa_mie {
  int nch = 0;
  std::vector<float> channels_m;
  // host staging (tracer-major), rebuilt on each add
  int nrh = 0, nmom = 0, npol = 0;
  std::vector<std::string> names;
  std::vector<int> nrh_t, nmom_t, npol_t;
  std::vector<float> bext, bsca, g, bbck, refr, refi, pmom;   // [tr][ch][rh], pmom [tr][ch][rh][pol][mom]
  std::vector<int> rhi;
  std::vector<float> rha;
  // per-tracer fields of tables built from the file arrays (aa_mie_add_table_native):
  // reff/gf/rhop/rhod/vol/area [rh], pback [ch][rh][pol]; empty for pre-interpolated tables
  std::vector<std::vector<float>> reff_t, gf_t, rhop_t, rhod_t, vol_t, area_t, pback_t;
  std::vector<int> npback_t;
  // table number and bin of every tracer (tables in the order they were added)
  std::vector<int> table_t, bin_t;
  int ntables = 0;
  // device copies, uploaded on first use after a change; `mu` serialises the
  // upload (and the table edits that mark it dirty) between host threads
  // that share one table set, e.g. the slots of a serving loop
  void* dmem[12] = {nullptr};
  aa::MieDev dev;
  bool dirty = true;
  std::mutex mu;
};

namespace {

std::string norm_name(const char* s, size_t len) {
  std::string n(s, len);
  size_t p = n.find("::");
  if (p != std::string::npos) n = n.substr(p + 2);
  while (!n.empty() && (n.back() == ' ' || n.back() == '\0')) n.pop_back();
  size_t b = 0;
  while (b < n.size() && n[b] == ' ') ++b;
  n = n.substr(b);
  for (auto& c : n) c = (char)std::toupper((unsigned char)c);
  return n;
}

// RH hash (CB/Chem_MieTableMod.F90:612-627), float32 arithmetic as the reference.
void build_rh_map(const float* rh, int nrh, int* rhi, float* rha) {
  for (int j = 0; j < aa::NRH_MAP; ++j) {
    rhi[j] = 0; rha[j] = 0.0f;
    for (int i = nrh - 1; i >= 0; --i) {
      volatile float prod = rh[i] * 1000.0f;
      int th = (int)prod;
      if (j >= th) {
        rhi[j] = i + 1;
        if (i + 1 >= nrh) rha[j] = 0.0f;
        else {
          volatile float a = (float)j / 1000.0f;
          volatile float num = a - rh[i];
          volatile float den = rh[i + 1] - rh[i];
          rha[j] = num / den;
        }
        break;
      }
    }
  }
}

void free_dev(aa_mie* t) {
  for (auto& p : t->dmem) { if (p) (void)hipFree(p); p = nullptr; }
}

int upload(aa_mie* t) {
  std::lock_guard<std::mutex> lk(t->mu);
  if (!t->dirty) return 0;
  // tables changed after an earlier upload: work of earlier calls may still
  // read the old copies on some stream
  bool had = false;
  for (void* p : t->dmem) had = had || p != nullptr;
  if (had) AA_CHECK(hipDeviceSynchronize());
  free_dev(t);
  auto put = [&](int slot, const void* src, size_t bytes) -> int {
    if (bytes == 0) bytes = 4;
    AA_CHECK(hipMalloc(&t->dmem[slot], bytes));
    AA_CHECK(hipMemcpy(t->dmem[slot], src, bytes, hipMemcpyHostToDevice));
    return 0;
  };
  int rc = 0;
  rc |= put(0, t->bext.data(), t->bext.size() * 4);
  rc |= put(1, t->bsca.data(), t->bsca.size() * 4);
  rc |= put(2, t->g.data(), t->g.size() * 4);
  rc |= put(3, t->bbck.data(), t->bbck.size() * 4);
  rc |= put(4, t->refr.data(), t->refr.size() * 4);
  rc |= put(5, t->refi.data(), t->refi.size() * 4);
  rc |= put(6, t->pmom.data(), t->pmom.size() * 4);
  rc |= put(7, t->rhi.data(), t->rhi.size() * 4);
  rc |= put(8, t->rha.data(), t->rha.size() * 4);
  rc |= put(9, t->nrh_t.data(), t->nrh_t.size() * 4);
  rc |= put(10, t->nmom_t.data(), t->nmom_t.size() * 4);
  rc |= put(11, t->npol_t.data(), t->npol_t.size() * 4);
  if (rc) return -2;
  aa::MieDev& d = t->dev;
  d.ntracer = (int)t->names.size(); d.nch = t->nch; d.nrh = t->nrh; d.nmom = t->nmom; d.npol = t->npol;
  d.bext = (const float*)t->dmem[0]; d.bsca = (const float*)t->dmem[1]; d.g = (const float*)t->dmem[2];
  d.bbck = (const float*)t->dmem[3]; d.refr = (const float*)t->dmem[4]; d.refi = (const float*)t->dmem[5];
  d.pmom = (const float*)t->dmem[6]; d.rhi = (const int*)t->dmem[7]; d.rha = (const float*)t->dmem[8];
  d.nrh_t = (const int*)t->dmem[9]; d.nmom_t = (const int*)t->dmem[10]; d.npol_t = (const int*)t->dmem[11];
  t->dirty = false;
  return 0;
}

}  // namespace

namespace aa {
// Device memory of the synchronous host-pointer entry points (vlidort.hip):
// one grow-only block per device, used under host_call_mutex.
std::mutex& host_call_mutex();
int host_call_block(size_t bytes, char** out);

// Accessors used by the fused radiance path.
int mie_upload(aa_mie* t) { return upload(t); }
// True when the device copies are current (no upload, hence no allocation,
// is needed before the tables are read).
bool mie_resident(aa_mie* t) {
  std::lock_guard<std::mutex> lk(t->mu);
  return !t->dirty;
}
const MieDev& mie_dev(const aa_mie* t) { return t->dev; }

int launch_get_aop(const aa_mie* t, int km, int nobs, int nch, int nq, const int* d_tracer_idx,
                   const int* d_chan_idx, const float* d_qm, const float* d_rh, int nmom, int npol, int ntr_eff,
                   int normalize, float* d_tau, float* d_ssa, float* d_g, float* d_pmom, hipStream_t st) {
  long ne = (long)km * nobs;
  if (ne == 0) return 0;
  const long blocks = (long)nobs * nch;
  if (ntr_eff > MAXTR || blocks > 0x7fffffffL) return -1;
  // table rows are addressed with 32-bit row indices
  if ((long)t->dev.ntracer * t->dev.nch * t->dev.nrh > 0x7fffffffL) return -1;
  prof_begin(PROF_MIE, st);
  hipLaunchKernelGGL(get_aop_kernel, dim3((unsigned)blocks), dim3(64), 0, st, t->dev, km, nobs, nch, nq,
                     d_tracer_idx, d_chan_idx, d_qm, d_rh, nmom, npol, ntr_eff, normalize, d_tau, d_ssa, d_g,
                     d_pmom);
  prof_end(PROF_MIE, st);
  AA_CHECK(hipGetLastError());
  return 0;
}
}  // namespace aa

extern "C" {

int aa_mie_create(aa_mie** out, int nch, const float* channels_m) {
  if (!out || nch <= 0 || !channels_m) return -1;
  aa_mie* t = new aa_mie();
  t->nch = nch;
  t->channels_m.assign(channels_m, channels_m + nch);
  *out = t;
  return 0;
}

static int add_table_locked(aa_mie* t, int nrh, int nbin, int nmom, int npol, const float* rh, const float* bext,
                            const float* bsca, const float* bbck, const float* g, const float* refr,
                            const float* refi, const float* pmom, const char* bin_names);

int aa_mie_add_table(aa_mie* t, int nrh, int nbin, int nmom, int npol, const float* rh, const float* bext,
                     const float* bsca, const float* bbck, const float* g, const float* refr, const float* refi,
                     const float* pmom, const char* bin_names) {
  if (!t) return -1;
  std::lock_guard<std::mutex> lk(t->mu);
  return add_table_locked(t, nrh, nbin, nmom, npol, rh, bext, bsca, bbck, g, refr, refi, pmom, bin_names);
}

// Body of aa_mie_add_table; the caller holds t->mu.
static int add_table_locked(aa_mie* t, int nrh, int nbin, int nmom, int npol, const float* rh, const float* bext,
                            const float* bsca, const float* bbck, const float* g, const float* refr,
                            const float* refi, const float* pmom, const char* bin_names) {
  if (!t || nrh <= 0 || nbin <= 0 || nmom < 0 || npol <= 0 || npol > 6 || !rh || !bext || !bsca || !g ||
      !bin_names)
    return -1;
  const int nch = t->nch;
  const int old_tr = (int)t->names.size();
  // common padded dimensions across tables
  int new_nrh = std::max(t->nrh, nrh), new_nmom = std::max(t->nmom, nmom), new_npol = std::max(t->npol, npol);
  auto regrid3 = [&](std::vector<float>& v) {   // [tr][ch][rh] with new nrh
    if (new_nrh == t->nrh) return;
    std::vector<float> nv((size_t)old_tr * nch * new_nrh, 0.0f);
    for (int tr = 0; tr < old_tr; ++tr) for (int c = 0; c < nch; ++c) for (int r = 0; r < t->nrh; ++r)
      nv[((size_t)tr * nch + c) * new_nrh + r] = v[((size_t)tr * nch + c) * t->nrh + r];
    v.swap(nv);
  };
  regrid3(t->bext); regrid3(t->bsca); regrid3(t->g); regrid3(t->bbck); regrid3(t->refr); regrid3(t->refi);
  if (new_nrh != t->nrh || new_nmom != t->nmom || new_npol != t->npol) {
    std::vector<float> nv((size_t)old_tr * nch * new_nrh * new_npol * new_nmom, 0.0f);
    for (int tr = 0; tr < old_tr; ++tr) for (int c = 0; c < nch; ++c) for (int r = 0; r < t->nrh; ++r)
      for (int p = 0; p < t->npol; ++p) for (int m = 0; m < t->nmom; ++m)
        nv[((((size_t)tr * nch + c) * new_nrh + r) * new_npol + p) * new_nmom + m] =
            t->pmom[((((size_t)tr * nch + c) * t->nrh + r) * t->npol + p) * t->nmom + m];
    t->pmom.swap(nv);
  }
  t->nrh = new_nrh; t->nmom = new_nmom; t->npol = new_npol;
  for (int b = 0; b < nbin; ++b) {
    t->names.push_back(norm_name(bin_names + 16 * b, 16));
    t->nrh_t.push_back(nrh); t->nmom_t.push_back(nmom); t->npol_t.push_back(npol);
    t->table_t.push_back(t->ntables); t->bin_t.push_back(b);
    for (int c = 0; c < nch; ++c) for (int r = 0; r < new_nrh; ++r) {
      size_t src = ((size_t)b * nch + c) * nrh + r;   // Fortran (nrh,nch,nbin)
      bool in = r < nrh;
      t->bext.push_back(in ? bext[src] : 0.0f);
      t->bsca.push_back(in ? bsca[src] : 0.0f);
      t->g.push_back(in ? g[src] : 0.0f);
      t->bbck.push_back(in && bbck ? bbck[src] : 0.0f);
      t->refr.push_back(in && refr ? refr[src] : 0.0f);
      t->refi.push_back(in && refi ? refi[src] : 0.0f);
    }
    for (int c = 0; c < nch; ++c) for (int r = 0; r < new_nrh; ++r) for (int p = 0; p < new_npol; ++p)
      for (int m = 0; m < new_nmom; ++m) {
        float v = 0.0f;
        if (pmom && r < nrh && p < npol && m < nmom)   // Fortran pmom(nrh,nch,nbin,nmom,npol)
          v = pmom[((((size_t)p * nmom + m) * nbin + b) * nch + c) * nrh + r];
        t->pmom.push_back(v);
      }
    std::vector<int> hi(aa::NRH_MAP);
    std::vector<float> ha(aa::NRH_MAP);
    build_rh_map(rh, nrh, hi.data(), ha.data());
    t->rhi.insert(t->rhi.end(), hi.begin(), hi.end());
    t->rha.insert(t->rha.end(), ha.begin(), ha.end());
    for (auto* v : {&t->reff_t, &t->gf_t, &t->rhop_t, &t->rhod_t, &t->vol_t, &t->area_t, &t->pback_t})
      v->emplace_back();
    t->npback_t.push_back(0);
  }
  ++t->ntables;
  t->dirty = true;
  return 0;
}

// Linear interpolation of Chem_MieTableRead's polint (CB/Chem_MieTableMod.F90:
// 671-717): float64 table abscissae x and values y, float32 target; slope
// and dx are default reals (float32), the product slope*dx is float32 and the
// sum with y(i) is float64, rounded to float32.  Below x(1) the first
// interval's slope is used from x(1); above x(n) the value is y(n).
static float polint_ref(const double* x, const double* y, int n, size_t ystride, float xw) {
  int i = 0;
  if ((double)xw < x[0]) i = 1;
  if ((double)xw > x[n - 1]) i = n;
  if (i == 0)
    for (int j = 1; j <= n; ++j)
      if ((double)xw >= x[j - 1]) i = j;
  if (i == 0) i = 1;   // unordered target (NaN): the reference leaves i = 0; clamp to stay in bounds
  float slope = 0.0f;
  if (i != n) slope = (float)((y[(size_t)i * ystride] - y[(size_t)(i - 1) * ystride]) / (x[i] - x[i - 1]));
  const float dx = (float)((double)xw - x[i - 1]);
  const float sd = slope * dx;
  return (float)(y[(size_t)(i - 1) * ystride] + (double)sd);
}

// Chem_MieTableRead (CB/Chem_MieTableMod.F90:238-719) from the arrays of an
// optics-table file: fill values of absent fields, derived volume and area,
// polint to the table set's channels, RH map; then the table is added as by
// aa_mie_add_table and the per-bin fields Chem_MieQuery can return are kept.
int aa_mie_add_table_native(aa_mie* t, int nch_table, const double* lambda_table, int nrh, int nbin, int nmom_table,
                            int npol_table, int nmom, const double* rh, const double* reff, const double* bext,
                            const double* bsca, const double* bbck, const double* g, const double* pback,
                            const double* pmom, const double* gf, const double* rhop, const double* refr,
                            const double* refi, const char* bin_names, int* rc) {
  if (!t || !rc || nch_table <= 0 || nrh <= 0 || nbin <= 0 || npol_table <= 0 || npol_table > 6 || nmom < 0 ||
      !lambda_table || !rh || !reff || !bext || !bsca || !bbck || !g || !bin_names)
    return -1;
  *rc = 0;
  if (nmom > 0 && nmom > nmom_table) { *rc = 99; return 0; }   // :296-301
  std::lock_guard<std::mutex> lk(t->mu);
  if (nmom > 0 && !pmom) return -1;
  const int nch = t->nch;
  const size_t nl = (size_t)nch_table, n3 = nl * nrh * nbin, n2 = (size_t)nrh * nbin;
  // fields of the file, with the reference's fill values where absent (:405-470)
  std::vector<double> pb(n3 * npol_table, 1.0), gfv(n2, -999.0), rhp(n2, -999.0), rhd(n2, -999.0);
  std::vector<double> rr(n3, -999.0), ri(n3, -999.0);
  if (pback) pb.assign(pback, pback + n3 * npol_table);
  if (gf) gfv.assign(gf, gf + n2);
  if (rhop) {
    rhp.assign(rhop, rhop + n2);
    // rhod is read from the same 'rhop' variable, first RH row broadcast (:433-444)
    for (int j = 0; j < nbin; ++j)
      for (int i = 0; i < nrh; ++i) rhd[(size_t)j * nrh + i] = rhop[(size_t)j * nrh];
  }
  if (refr) rr.assign(refr, refr + n3);
  if (refi) for (size_t k = 0; k < n3; ++k) ri[k] = std::fabs(refi[k]);
  // wet volume and cross-section area (:487-491)
  std::vector<double> vol(n2), area(n2);
  for (size_t k = 0; k < n2; ++k) {
    vol[k] = gfv[k] * gfv[k] * gfv[k] / rhd[k];
    area[k] = vol[k] / ((double)(4.0f / 3.0f) * reff[k]);
  }
  // polint to the channels (:560-589); outputs in aa_mie_add_table's layouts
  const size_t o3 = (size_t)nrh * nch * nbin;
  std::vector<float> ob(o3), os(o3), obb(o3), og(o3), orr(o3), ori(o3), opm(o3 * (size_t)std::max(nmom, 0) * npol_table);
  std::vector<float> opb(o3 * npol_table), rhf(nrh);
  for (int i = 0; i < nrh; ++i) rhf[i] = (float)rh[i];
  for (int j = 0; j < nbin; ++j)
    for (int i = 0; i < nrh; ++i)
      for (int c = 0; c < nch; ++c) {
        const float xw = t->channels_m[c];
        const size_t src = ((size_t)j * nrh + i) * nl;        // (nch_table, nrh, nbin)
        const size_t dst = ((size_t)j * nch + c) * nrh + i;   // (nrh, nch, nbin)
        ob[dst] = polint_ref(lambda_table, bext + src, nch_table, 1, xw);
        os[dst] = polint_ref(lambda_table, bsca + src, nch_table, 1, xw);
        obb[dst] = polint_ref(lambda_table, bbck + src, nch_table, 1, xw);
        og[dst] = polint_ref(lambda_table, g + src, nch_table, 1, xw);
        orr[dst] = polint_ref(lambda_table, rr.data() + src, nch_table, 1, xw);
        ori[dst] = polint_ref(lambda_table, ri.data() + src, nch_table, 1, xw);
        for (int p = 0; p < npol_table; ++p)
          opb[dst + o3 * p] = polint_ref(lambda_table, pb.data() + src + n3 * p, nch_table, 1, xw);
        for (int m = 0; m < nmom; ++m)
          for (int p = 0; p < npol_table; ++p)   // file pmom(nch_table, nrh, nbin, nmom_table, npol)
            opm[dst + o3 * ((size_t)p * nmom + m)] =
                polint_ref(lambda_table, pmom + src + n3 * ((size_t)p * nmom_table + m), nch_table, 1, xw);
      }
  const int old_tr = (int)t->names.size();
  int e = add_table_locked(t, nrh, nbin, nmom, npol_table, rhf.data(), ob.data(), os.data(), obb.data(), og.data(),
                           orr.data(), ori.data(), nmom > 0 ? opm.data() : nullptr, bin_names);
  if (e) return e;
  for (int j = 0; j < nbin; ++j) {
    const int tr = old_tr + j;
    auto f32 = [&](const double* a) {
      std::vector<float> v(nrh);
      for (int i = 0; i < nrh; ++i) v[i] = (float)a[(size_t)j * nrh + i];
      return v;
    };
    t->reff_t[tr] = f32(reff); t->gf_t[tr] = f32(gfv.data()); t->rhop_t[tr] = f32(rhp.data());
    t->rhod_t[tr] = f32(rhd.data()); t->vol_t[tr] = f32(vol.data()); t->area_t[tr] = f32(area.data());
    std::vector<float> pbt((size_t)nch * nrh * npol_table);
    for (int c = 0; c < nch; ++c)
      for (int i = 0; i < nrh; ++i)
        for (int p = 0; p < npol_table; ++p)
          pbt[((size_t)c * nrh + i) * npol_table + p] = opb[((size_t)j * nch + c) * nrh + i + o3 * p];
    t->pback_t[tr] = std::move(pbt);
    t->npback_t[tr] = npol_table;
  }
  return 0;
}

// The Chem_MieQuery outputs beyond the optics (CB/Chem_MieMod.F90:885-943):
// rEff [um], p11, p22 (pback polarisations 1 and 5), growth factor, dry
// density (first RH), wet volume and area, interpolated in RH with the same
// float32 arithmetic as the optics.  Only for tables built by
// aa_mie_add_table_native.  (The reference's rhop output is never assigned
// there, so it is not offered.)
int aa_mie_query_extra(const aa_mie* t, int idx, float channel, int n, const float* rh, float* reff, float* p11,
                       float* p22, float* gf, float* rhod, float* vol, float* area) {
  if (!t || idx < 0 || idx >= (int)t->names.size() || n < 0 || !rh) return -1;
  if (t->reff_t[idx].empty()) return -1;
  const int ich = (int)lrintf(channel) - 1;
  if (ich < 0 || ich >= t->nch) return -1;
  const int np = t->npback_t[idx], nrh = t->nrh_t[idx];
  if ((p11 || p22) && np < 5) return -1;
  const int* hi = t->rhi.data() + (size_t)idx * aa::NRH_MAP;
  const float* ha = t->rha.data() + (size_t)idx * aa::NRH_MAP;
  const std::vector<float>& pb = t->pback_t[idx];
  for (int k = 0; k < n; ++k) {
    float rhu = std::min(rh[k], 0.99f);
    int isnap = (int)((rhu + 0.001f) * 1000.0f);
    if (isnap < 1) isnap = 1;
    const float arh = ha[isnap - 1];
    const int irh = hi[isnap - 1] - 1;
    const int irhp1 = std::min(irh + 1, nrh - 1);
    const float w0 = 1.0f - arh;
    auto lin = [&](const std::vector<float>& v) {
      const float a = v[irh] * w0;
      const float b = v[irhp1] * arh;
      return a + b;
    };
    if (reff) { const float r = lin(t->reff_t[idx]); reff[k] = 1.0e6f * r; }
    if (p11) {
      const float a = pb[((size_t)ich * nrh + irh) * np + 0] * w0;
      const float b = pb[((size_t)ich * nrh + irhp1) * np + 0] * arh;
      p11[k] = a + b;
    }
    if (p22) {
      const float a = pb[((size_t)ich * nrh + irh) * np + 4] * w0;
      const float b = pb[((size_t)ich * nrh + irhp1) * np + 4] * arh;
      p22[k] = a + b;
    }
    if (gf) gf[k] = lin(t->gf_t[idx]);
    if (rhod) rhod[k] = t->rhod_t[idx][0];
    if (vol) vol[k] = lin(t->vol_t[idx]);
    if (area) area[k] = lin(t->area_t[idx]);
  }
  return 0;
}

int aa_mie_query_idx(const aa_mie* t, const char* qname) {
  if (!t || !qname) return -1;
  std::string n = norm_name(qname, std::strlen(qname));
  if (n.empty()) return -1;   // bins no tracer maps to have empty names
  for (size_t i = 0; i < t->names.size(); ++i) if (t->names[i] == n) return (int)i;
  return -1;
}

int aa_mie_channels(const aa_mie* t, float* channels_m) {
  if (!t || !channels_m) return -1;
  for (int c = 0; c < t->nch; ++c) channels_m[c] = t->channels_m[c];
  return 0;
}

int aa_mie_tracer_table(const aa_mie* t, int idx, int* table, int* bin, int* nbin, int* nrh, int* nmom,
                        int* npol) {
  if (!t || idx < 0 || idx >= (int)t->names.size()) return -1;
  const int tb = t->table_t[idx];
  int nb = 0;
  for (int v : t->table_t) nb += (v == tb);
  if (table) *table = tb;
  if (bin) *bin = t->bin_t[idx];
  if (nbin) *nbin = nb;
  if (nrh) *nrh = t->nrh_t[idx];
  if (nmom) *nmom = t->nmom_t[idx];
  if (npol) *npol = t->npol_t[idx];
  return 0;
}

int aa_mie_dims(const aa_mie* t, int* nch, int* nmom_min, int* npol_max, int* ntracer) {
  if (!t) return -1;
  int mm = 1 << 30;
  for (int v : t->nmom_t) mm = std::min(mm, v);
  if (nch) *nch = t->nch;
  if (nmom_min) *nmom_min = t->names.empty() ? 0 : mm;
  if (npol_max) *npol_max = t->npol;
  if (ntracer) *ntracer = (int)t->names.size();
  return 0;
}

int aa_mie_query(const aa_mie* tc, int idx, float channel, int n, const float* q_mass, const float* rh, float* tau,
                 float* ssa, float* gasym, float* bext, float* bsca, float* bbck, float* refr, float* refi,
                 int nmom, int npol, float* pmom) {
  aa_mie* t = const_cast<aa_mie*>(tc);
  if (!t || idx < 0 || idx >= (int)t->names.size() || n < 0 || !q_mass || !rh) return -1;
  int ich = (int)lrintf(channel) - 1;
  if (ich < 0 || ich >= t->nch) return -1;
  if (pmom && (nmom > t->nmom_t[idx] || npol > t->npol_t[idx])) return -1;
  if (n == 0) return 0;
  if (upload(t)) return -2;
  float* outs[9] = {tau, ssa, gasym, bext, bsca, bbck, refr, refi, pmom};
  size_t bytes[9];
  for (int i = 0; i < 8; ++i) bytes[i] = (size_t)n * 4;
  bytes[8] = (size_t)n * (pmom ? nmom * npol : 0) * 4;
  std::lock_guard<std::mutex> hl(aa::host_call_mutex());
  // device copies carved from the host-call block, 256-byte aligned
  size_t off[12], tot = 0;
  auto carve = [&](int k, size_t b) { off[k] = tot; tot += (b + 255) & ~(size_t)255; };
  carve(0, (size_t)n * 4);
  carve(1, (size_t)n * 4);
  for (int i = 0; i < 9; ++i) carve(2 + i, (outs[i] && bytes[i]) ? bytes[i] : 0);
  char* base = nullptr;
  if (aa::host_call_block(tot, &base)) return -2;
  void* dq = base + off[0];
  void* drh = base + off[1];
  void* dout[9] = {nullptr};
  for (int i = 0; i < 9; ++i) if (outs[i] && bytes[i]) dout[i] = base + off[2 + i];
  AA_CHECK(hipMemcpy(dq, q_mass, (size_t)n * 4, hipMemcpyHostToDevice));
  AA_CHECK(hipMemcpy(drh, rh, (size_t)n * 4, hipMemcpyHostToDevice));
  int threads = 256;
  hipLaunchKernelGGL(aa::mie_query_kernel, dim3((n + threads - 1) / threads), dim3(threads), 0, 0, t->dev, idx, ich,
                     n, (const float*)dq, (const float*)drh, (float*)dout[0], (float*)dout[1], (float*)dout[2],
                     (float*)dout[3], (float*)dout[4], (float*)dout[5], (float*)dout[6], (float*)dout[7], nmom, npol,
                     (float*)dout[8]);
  AA_CHECK(hipGetLastError());
  AA_CHECK(hipDeviceSynchronize());
  for (int i = 0; i < 9; ++i) if (dout[i]) AA_CHECK(hipMemcpy(outs[i], dout[i], bytes[i], hipMemcpyDeviceToHost));
  return 0;
}

int aa_mie_destroy(aa_mie* t) {
  if (!t) return -1;
  free_dev(t);
  delete t;
  return 0;
}

// Host-pointer VLIDORT_getAOPvector.
static int get_aop_host(aa_mie* t, int km, int nobs, int nch, int nq, const float* channels, const char* vname,
                        int verbose, const float* qm, const float* rh, int nmom, int npol, float* tau, float* ssa,
                        float* g, float* pmom, int* rc) {
  if (!t || km <= 0 || nobs < 0 || nch <= 0 || nq <= 0 || nq > aa::MAXTR || !channels || !vname || !qm || !rh ||
      !tau || !ssa || !g || !rc || npol > 6)
    return -1;
  *rc = 0;
  const size_t ne = (size_t)km * nch * nobs;
  // outputs start at zero (GEO/VLIDORT_MieMod.F90:234-239); the kernel writes
  // every element, so only the channel-mismatch return clears them here (the
  // reference returns rc 99 before its zeroing and leaves them undefined)
  auto clear = [&]() {
    std::fill(tau, tau + ne, 0.0f); std::fill(ssa, ssa + ne, 0.0f); std::fill(g, g + ne, 0.0f);
    if (pmom) std::fill(pmom, pmom + ne * nmom * npol, 0.0f);
  };
  // channel matching in float32 (GEO/VLIDORT_MieMod.F90:210-225)
  std::vector<int> chan(nch, -1);
  for (int n = 0; n < nch; ++n)
    for (int m = 0; m < t->nch; ++m) {
      volatile float scaled = 1.0e9f * t->channels_m[m];
      volatile float diff = channels[n] - scaled;
      if (std::fabs(diff) < 1.0f) { chan[n] = m; break; }
    }
  for (int n = 0; n < nch; ++n) if (chan[n] < 0) { clear(); *rc = 99; return 0; }
  // tracer resolution and the reference's early returns
  std::vector<int> tidx(nq, -1);
  int ntr_eff = nq, normalize = 1, rcv = 0;
  for (int iq = 0; iq < nq; ++iq) {
    char nm[17];
    for (int c = 0; c < 16; ++c) nm[c] = vname[(size_t)c * nq + iq];
    nm[16] = 0;
    int id = aa_mie_query_idx(t, nm);
    rcv = (id < 0) ? 1 : 0;
    if (id < 0) continue;
    if (verbose == 1) std::printf("[+] Adding %s contribution\n", t->names[id].c_str());
    if (nmom > t->nmom_t[id]) { ntr_eff = iq; normalize = 0; rcv = 666; break; }
    int tp = t->npol_t[id];
    if (!(npol <= tp || (npol == 6 && tp == 4))) { ntr_eff = iq; normalize = 0; rcv = 777; break; }
    tidx[iq] = id;
  }
  if (nobs == 0) { *rc = rcv; return 0; }
  if (upload(t)) return -2;
  std::lock_guard<std::mutex> hl(aa::host_call_mutex());
  // device copies carved from the host-call block, 256-byte aligned
  const size_t sz[8] = {(size_t)km * nq * nobs * 4, (size_t)km * nobs * 4, ne * 4, ne * 4, ne * 4,
                        (pmom && nmom > 0) ? ne * nmom * npol * 4 : 0, (size_t)nq * 4, (size_t)nch * 4};
  size_t off[8], tot = 0;
  for (int i = 0; i < 8; ++i) { off[i] = tot; tot += (sz[i] + 255) & ~(size_t)255; }
  char* base = nullptr;
  if (aa::host_call_block(tot, &base)) return -2;
  void *dq = base + off[0], *drh = base + off[1], *dt = base + off[2], *ds = base + off[3], *dg = base + off[4];
  void *dp = sz[5] ? base + off[5] : nullptr, *dti = base + off[6], *dci = base + off[7];
  AA_CHECK(hipMemcpy(dq, qm, (size_t)km * nq * nobs * 4, hipMemcpyHostToDevice));
  AA_CHECK(hipMemcpy(drh, rh, (size_t)km * nobs * 4, hipMemcpyHostToDevice));
  AA_CHECK(hipMemcpy(dti, tidx.data(), nq * 4, hipMemcpyHostToDevice));
  AA_CHECK(hipMemcpy(dci, chan.data(), nch * 4, hipMemcpyHostToDevice));
  int e = aa::launch_get_aop(t, km, nobs, nch, nq, (const int*)dti, (const int*)dci, (const float*)dq,
                             (const float*)drh, nmom, npol, ntr_eff, normalize, (float*)dt, (float*)ds, (float*)dg,
                             (float*)dp, 0);
  if (e) return e;
  AA_CHECK(hipDeviceSynchronize());
  AA_CHECK(hipMemcpy(tau, dt, ne * 4, hipMemcpyDeviceToHost));
  AA_CHECK(hipMemcpy(ssa, ds, ne * 4, hipMemcpyDeviceToHost));
  AA_CHECK(hipMemcpy(g, dg, ne * 4, hipMemcpyDeviceToHost));
  if (dp) AA_CHECK(hipMemcpy(pmom, dp, ne * nmom * npol * 4, hipMemcpyDeviceToHost));
  *rc = rcv;
  return 0;
}

int aa_get_aop_vector(aa_mie* t, int km, int nobs, int nch, int nq, const float* channels, const char* vname,
                      int verbose, const float* qm, const float* rh, int nmom, int npol, float* tau, float* ssa,
                      float* g, float* pmom, int* rc) {
  if (!pmom) return -1;
  return get_aop_host(t, km, nobs, nch, nq, channels, vname, verbose, qm, rh, nmom, npol, tau, ssa, g, pmom, rc);
}

int aa_get_aop_scalar(aa_mie* t, int km, int nobs, int nch, int nq, const float* channels, const char* vname,
                      int verbose, const float* qm, const float* rh, float* tau, float* ssa, float* g, int* rc) {
  return get_aop_host(t, km, nobs, nch, nq, channels, vname, verbose, qm, rh, 0, 1, tau, ssa, g, nullptr, rc);
}

int aa_get_aop_vector_dev(aa_mie* t, int km, int nobs, int nch, int nq, const int* tracer_idx, const int* chan_idx,
                          const float* d_qm, const float* d_rh, int nmom, int npol, float* d_tau, float* d_ssa,
                          float* d_g, float* d_pmom, void* stream) {
  if (!t || km <= 0 || nobs < 0 || nch <= 0 || nq <= 0 || nq > aa::MAXTR || npol > 6) return -1;
  if (upload(t)) return -2;
  return aa::launch_get_aop(t, km, nobs, nch, nq, tracer_idx, chan_idx, d_qm, d_rh, nmom, npol, nq, 1, d_tau, d_ssa,
                            d_g, d_pmom, (hipStream_t)stream);
}

}  // extern "C"
