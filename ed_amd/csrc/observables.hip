// observables.hip -- correlations, local magnetisation, S^a_q |psi> and the
// DSF continued-fraction recurrence, on the device.
#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

#include "blas.h"
#include "qs3x_internal.h"
#include "rep.h"

using namespace qs3x;

namespace {

constexpr int kObsThreads = 256;

template <class F>
int guarded3(F &&f) {
  try {
    f();
    return 0;
  } catch (const Error &e) {
    set_error(e.msg);
    return 1;
  } catch (const std::exception &e) {
    set_error(std::string("qs3x: ") + e.what());
    return 2;
  }
}

__device__ void block_sum3(double a, double b, double c, double *out) {
  __shared__ double red[3][kObsThreads / 64];
  for (int d = 32; d >= 1; d >>= 1) {
    a += __shfl_down(a, d, 64);
    b += __shfl_down(b, d, 64);
    c += __shfl_down(c, d, 64);
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wid] = a;
    red[1][wid] = b;
    red[2][wid] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = 0, y = 0, z = 0;
    for (int w = 0; w < kObsThreads / 64; ++w) {
      x += red[0][w];
      y += red[1][w];
      z += red[2][w];
    }
    out[0] = x;
    out[1] = y;
    out[2] = z;
  }
  __syncthreads();
}

// calcu_cf_trans_2 (ham2vec_v3.f90:111-174) for many pair orbits at once.
// blockIdx.y = orbit; pairs[2*(off[o] + q) + {0,1}] = (list_ij(1,q), list_ij(2,q)).
// part[3*(o*nbx + bx) + {0,1,2}] = (szsz, Re spsm, Im spsm) block sums.
__global__ __launch_bounds__(kObsThreads) void k_corr(SecView v, const double2 *__restrict__ psi,
                                                      const int32_t *__restrict__ off,
                                                      const int32_t *__restrict__ pairs, double *__restrict__ part) {
  const int o = blockIdx.y;
  const int q0 = off[o], q1 = off[o + 1];
  const bool diag = pairs[2 * q0] == pairs[2 * q0 + 1];
  double sz = 0, spr = 0, spi = 0;
  int n[kMaxNod], nn[kMaxNod];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < v.dim; i += (int64_t)gridDim.x * blockDim.x) {
    sv_unpack(v, i, n);
    const double2 x = psi[i];
    const double wgt = x.x * x.x + x.y * x.y;
    const double ri = sv_rowr(v, i);
    for (int q = q0; q < q1; ++q) {
      const int a = pairs[2 * q], b = pairs[2 * q + 1];
      bool f1 = false, f2 = false;
      for (int t = 0; t < v.N; ++t) {
        f1 |= (n[t] == a);
        f2 |= (n[t] == b);
      }
      if (diag) {
        if (!f1) spr += wgt;
        continue;
      }
      if (!f1 && f2) {
        sz -= wgt;
      } else if (f1 && !f2) {
        sz -= wgt;
        sv_move(n, v.N, a, b, nn);  // j_flip_ni(min, max): the down spin moves from a to b
        int ei = 0, isrep = 0;
        double rt = 0;
        const int64_t id = sv_lookup(v, nn, &ei, &isrep, &rt);
        if (id >= 0) {
          const double c = rt / ri;
          const double er = v.explist[2 * ei] * c, eim = v.explist[2 * ei + 1] * c;
          const double2 y = psi[id];
          // e * conj(x) * y
          const double cr = x.x * y.x + x.y * y.y, ci = x.x * y.y - x.y * y.x;
          spr += er * cr - eim * ci;
          spi += er * ci + eim * cr;
        }
      } else {
        sz += wgt;
      }
    }
  }
  double out[3];
  __shared__ double o3[3];
  block_sum3(sz, spr, spi, o3);
  if (threadIdx.x == 0) {
    double *p = part + 3 * ((size_t)o * gridDim.x + blockIdx.x);
    p[0] = o3[0];
    p[1] = o3[1];
    p[2] = o3[2];
  }
  (void)out;
}

// calcu_lm_trans_2 (ham2vec_v3.f90:74-109): per site orbit, sum_rows |v|^2 sum_j (+-1).
__global__ __launch_bounds__(kObsThreads) void k_lm(SecView v, const double2 *__restrict__ psi,
                                                    const int32_t *__restrict__ off, const int32_t *__restrict__ sites,
                                                    double *__restrict__ part) {
  const int o = blockIdx.y;
  double acc = 0;
  int n[kMaxNod];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < v.dim; i += (int64_t)gridDim.x * blockDim.x) {
    sv_unpack(v, i, n);
    const double2 x = psi[i];
    const double wgt = x.x * x.x + x.y * x.y;
    for (int q = off[o]; q < off[o + 1]; ++q) {
      bool f = false;
      for (int t = 0; t < v.N; ++t) f |= (n[t] == sites[q]);
      acc += f ? -wgt : wgt;
    }
  }
  __shared__ double o3[3];
  block_sum3(acc, 0.0, 0.0, o3);
  if (threadIdx.x == 0) part[(size_t)o * gridDim.x + blockIdx.x] = o3[0];
}

// S^a_q |psi> in pull form (calcu_Sq_a, get_expectation_values.f90:281-377).
// For each target row (configuration m, the representative of its orbit) the
// sources are the pairs (GS row l, site i) whose image n_l (+/-) i lies in the
// orbit of m.  Enumerating the distinct translates y = T^h m and the sites i
// with y (-/+) i a GS representative visits each source exactly once; the
// weight is psi(l) r_new(lp)/r(l) e^{-iq.R_i} conj(explist_new(ell(y))).
// Chains with spsmsz = 2 only visit the translates that contain site 1 (a GS
// representative always does); other cases scan all translates.
__device__ void sq_accumulate(const SecView &gs, const SecView &tg, int spsmsz, const int *y, double rt_lp,
                              const double2 *__restrict__ psi, const double2 *__restrict__ expks, double &accr,
                              double &acci) {
  const int Nt = tg.N;
  int src[kMaxNod];
  // ell of y -> m in the target sector; weight conj(explist_new(ell))
  int ei = 0, isrep = 0;
  double rt = 0;
  (void)sv_lookup(tg, y, &ei, &isrep, &rt);
  const double er = tg.explist[2 * ei], eim = -tg.explist[2 * ei + 1];
  auto add = [&](int64_t l, double wr, double wi) {
    const double c = rt_lp / gs.list_r[l];
    const double2 ps = psi[l];
    const double ar = c * (ps.x * wr - ps.y * wi), ai = c * (ps.x * wi + ps.y * wr);
    accr += ar * er - ai * eim;
    acci += ar * eim + ai * er;
  };
  if (spsmsz == 3) {  // source n_l = y, weight sum_i Siz(i) e^{-iq R_i}
    int gi = 0, grep = 0;
    double grt = 0;
    const int64_t l = sv_lookup(gs, y, &gi, &grep, &grt);
    if (l < 0 || !grep) return;
    double sr = 0, si = 0;
    for (int i = 1; i <= gs.L; ++i) {
      bool dn = false;
      for (int q = 0; q < Nt; ++q) dn |= (y[q] == i);
      const double siz = dn ? -0.5 : 0.5;
      sr += siz * expks[i - 1].x;
      si += siz * expks[i - 1].y;
    }
    add(l, sr, si);
  } else if (spsmsz == 2) {  // sources: i in y, n_l = y \ {i}
    for (int r = 0; r < Nt; ++r) {
      const int i = y[r];
      int p = 0;
      for (int q = 0; q < Nt; ++q)
        if (q != r) src[p++] = y[q];
      int gi = 0, grep = 0;
      double grt = 0;
      const int64_t l = sv_lookup(gs, src, &gi, &grep, &grt);
      if (l < 0 || !grep) continue;
      add(l, expks[i - 1].x, expks[i - 1].y);
    }
  } else {  // spsmsz == 1: sources i not in y, n_l = y + {i}
    for (int i = 1; i <= gs.L; ++i) {
      bool in = false;
      for (int q = 0; q < Nt; ++q) in |= (y[q] == i);
      if (in) continue;
      int p = 0;
      bool put = false;
      for (int q = 0; q < Nt; ++q) {
        if (!put && y[q] > i) {
          src[p++] = i;
          put = true;
        }
        src[p++] = y[q];
      }
      if (!put) src[p++] = i;
      int gi = 0, grep = 0;
      double grt = 0;
      const int64_t l = sv_lookup(gs, src, &gi, &grep, &grt);
      if (l < 0 || !grep) continue;
      add(l, expks[i - 1].x, expks[i - 1].y);
    }
  }
}

__device__ __forceinline__ bool same_set(const int *a, const int *b, int n) {
  for (int q = 0; q < n; ++q)
    if (a[q] != b[q]) return false;
  return true;
}

__global__ __launch_bounds__(kObsThreads) void k_sq(SecView gs, SecView tg, int spsmsz,
                                                    const double2 *__restrict__ psi,
                                                    const double2 *__restrict__ expks, double2 *__restrict__ f,
                                                    double norm) {
  const int Nt = tg.N;
  int m[kMaxNod], y[kMaxNod], yprev[kMaxNod], tmp[kMaxNod];
  const bool chain_fast = gs.chain && tg.chain && spsmsz == 2;
  for (int64_t lp = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; lp < tg.dim;
       lp += (int64_t)gridDim.x * blockDim.x) {
    sv_unpack(tg, lp, m);
    const double rt_lp = tg.list_r[lp];
    double accr = 0, acci = 0;
    if (spsmsz == 3) {
      // the only GS representative in the orbit of m is m itself
      sq_accumulate(gs, tg, spsmsz, m, rt_lp, psi, expks, accr, acci);
    } else if (chain_fast) {
      // a GS representative contains site 1, so y = m + i does too: only the
      // rotations of m that bring one of its down spins to site 1 qualify
      const int L = tg.L;
      for (int j = 0; j < Nt; ++j) {
        for (int q = 0; q < Nt; ++q) {
          const int s = m[(j + q) % Nt] - m[j];
          y[q] = 1 + (s < 0 ? s + L : s);
        }
        bool seen = false;
        for (int j2 = 0; j2 < j && !seen; ++j2) {
          for (int q = 0; q < Nt; ++q) {
            const int s = m[(j2 + q) % Nt] - m[j2];
            yprev[q] = 1 + (s < 0 ? s + L : s);
          }
          seen = same_set(y, yprev, Nt);
        }
        if (!seen) sq_accumulate(gs, tg, spsmsz, y, rt_lp, psi, expks, accr, acci);
      }
    } else {
      // every distinct translate of m
      const int G = tg.LX * tg.LY * tg.LZ;
      for (int h = 0; h < G; ++h) {
        const int hi = h % tg.LX, hj = (h / tg.LX) % tg.LY, hk = h / (tg.LX * tg.LY);
        for (int q = 0; q < Nt; ++q) y[q] = sv_shift_site(m[q], hi, hj, hk, tg);
        for (int q = 1; q < Nt; ++q) {
          int t = y[q], z = q - 1;
          while (z >= 0 && y[z] > t) {
            y[z + 1] = y[z];
            --z;
          }
          y[z + 1] = t;
        }
        bool seen = false;
        for (int h2 = 0; h2 < h && !seen; ++h2) {
          const int i2 = h2 % tg.LX, j2 = (h2 / tg.LX) % tg.LY, k2 = h2 / (tg.LX * tg.LY);
          for (int q = 0; q < Nt; ++q) tmp[q] = sv_shift_site(m[q], i2, j2, k2, tg);
          for (int q = 1; q < Nt; ++q) {
            int t = tmp[q], z = q - 1;
            while (z >= 0 && tmp[z] > t) {
              tmp[z + 1] = tmp[z];
              --z;
            }
            tmp[z + 1] = t;
          }
          seen = same_set(y, tmp, Nt);
        }
        if (!seen) sq_accumulate(gs, tg, spsmsz, y, rt_lp, psi, expks, accr, acci);
      }
    }
    f[lp] = make_double2(accr * norm, acci * norm);
  }
}

// mk_list_ij (get_expectation_values.f90:97-131) on the host.
void mk_list_ij(const qs3x_model *m, int r0, int r1, std::vector<int> &lst, int &noij) {
  auto sh = [&](int s, int dx, int dy, int dz) {
    int s0 = s - 1;
    int x = s0 % m->LX, y = (s0 / m->LX) % m->LY, z = s0 / (m->LX * m->LY);
    x = (x + dx) % m->LX;
    y = (y + dy) % m->LY;
    z = (z + dz) % m->LZ;
    return 1 + x + m->LX * (y + m->LY * z);
  };
  lst.assign({r0, r1});
  noij = 1;
  int rd0 = r0, rd1 = r1;
  for (int k = 1; k <= m->LZ; ++k) {
    rd0 = sh(rd0, 0, 0, 1);
    rd1 = sh(rd1, 0, 0, 1);
    for (int j = 1; j <= m->LY; ++j) {
      rd0 = sh(rd0, 0, 1, 0);
      rd1 = sh(rd1, 0, 1, 0);
      for (int i = 1; i <= m->LX; ++i) {
        rd0 = sh(rd0, 1, 0, 0);
        rd1 = sh(rd1, 1, 0, 0);
        const int rd[2] = {rd0, rd1}, r[2] = {r0, r1};
        int l;
        for (l = 2; l >= 1; --l) {
          if (rd[l - 1] < r[l - 1]) {
            noij = 0;
            return;
          } else if (rd[l - 1] > r[l - 1]) {
            lst.push_back(rd0);
            lst.push_back(rd1);
            ++noij;
            break;
          }
        }
        if (l == 0) return;
      }
    }
  }
}

// mk_list_n (get_expectation_values.f90:65-95) on the host.
void mk_list_n(const qs3x_model *m, int n, std::vector<int> &lst) {
  auto sh = [&](int s, int dx, int dy, int dz) {
    int s0 = s - 1;
    int x = s0 % m->LX, y = (s0 / m->LX) % m->LY, z = s0 / (m->LX * m->LY);
    x = (x + dx) % m->LX;
    y = (y + dy) % m->LY;
    z = (z + dz) % m->LZ;
    return 1 + x + m->LX * (y + m->LY * z);
  };
  lst.assign({n});
  int nd = n;
  for (int k = 1; k <= m->LZ; ++k) {
    nd = sh(nd, 0, 0, 1);
    for (int j = 1; j <= m->LY; ++j) {
      nd = sh(nd, 0, 1, 0);
      for (int i = 1; i <= m->LX; ++i) {
        nd = sh(nd, 1, 0, 0);
        if (nd < n) {
          lst.clear();
          return;
        } else if (nd > n) {
          lst.push_back(nd);
        } else {
          return;
        }
      }
    }
  }
}

int obs_grid(int64_t dim) {
  int64_t b = (dim + kObsThreads - 1) / kObsThreads;
  return (int)std::max<int64_t>(1, std::min<int64_t>(b, 512));
}

}  // namespace

extern "C" {

int qs3x_correlations(qs3x_sector *s, const double *psi, int64_t npairs, const int32_t *pairs, double *szsz,
                      double *spsm) {
  return guarded3([&] {
    QS3X_REQUIRE(!is_u1(s), "qs3x_correlations: not available for U(1)-only sectors");
    QS3X_REQUIRE(s && psi && pairs && szsz && spsm, "qs3x_correlations: NULL argument");
    qs3x_model *m = s->model;
    const int NOS = m->NOS;
    // get_cf_2_wave_vector: list_expe_ss(1,j) = file column 2, (2,j) = column 1
    std::vector<int> la(npairs), lb(npairs);
    for (int64_t j = 0; j < npairs; ++j) {
      la[j] = pairs[2 * j + 1];
      lb[j] = pairs[2 * j];
      QS3X_REQUIRE(la[j] >= 1 && la[j] <= NOS && lb[j] >= 1 && lb[j] <= NOS, "qs3x_correlations: site out of range");
    }
    std::vector<int32_t> off{0}, flat;
    std::vector<std::vector<int>> orbits;
    const int64_t jmax = (int64_t)NOS * NOS;
    int64_t no_tot = 0;
    for (int64_t j = 0; j < std::min(jmax, npairs); ++j) {
      if (no_tot == jmax) break;
      std::vector<int> lst;
      int noij = 0;
      mk_list_ij(m, la[j], lb[j], lst, noij);
      if (noij == 0) continue;
      no_tot += noij;
      orbits.push_back(lst);
      for (int v : lst) flat.push_back(v);
      off.push_back((int32_t)(flat.size() / 2));
    }
    const int norb = (int)orbits.size();
    std::vector<double> ez((size_t)NOS * NOS, 0.0), esr((size_t)NOS * NOS, 0.0), esi((size_t)NOS * NOS, 0.0);
    if (norb > 0 && s->dim > 0) {
      hipStream_t st = stream();
      DevBuf<double> dpsi((size_t)2 * s->dim);
      dpsi.upload(psi, (size_t)2 * s->dim);
      DevBuf<int32_t> doff(off.size()), dpairs(std::max<size_t>(flat.size(), 1));
      doff.upload(off.data(), off.size());
      dpairs.upload(flat.data(), flat.size());
      const int gx = obs_grid(s->dim);
      DevBuf<double> part((size_t)3 * gx * norb);
      hipLaunchKernelGGL(k_corr, dim3(gx, norb), dim3(kObsThreads), 0, st, make_view(s), (const double2 *)dpsi.p,
                         doff.p, dpairs.p, part.p);
      QS3X_HIP(hipGetLastError());
      std::vector<double> hp(part.n);
      part.download(hp.data(), hp.size());
      for (int o = 0; o < norb; ++o) {
        double a = 0, b = 0, c = 0;
        for (int x = 0; x < gx; ++x) {
          a += hp[3 * ((size_t)o * gx + x)];
          b += hp[3 * ((size_t)o * gx + x) + 1];
          c += hp[3 * ((size_t)o * gx + x) + 2];
        }
        const int noij = (int)orbits[o].size() / 2;
        double vz;
        if (orbits[o][0] == orbits[o][1]) vz = 0.25;
        else vz = a / noij * 0.25;
        const double vr = b / noij, vi = c / noij;
        for (int q = 0; q < noij; ++q) {
          const size_t idx = (size_t)(orbits[o][2 * q] - 1) + (size_t)(orbits[o][2 * q + 1] - 1) * NOS;
          ez[idx] = vz;
          esr[idx] = vr;
          esi[idx] = vi;
        }
      }
    }
    // output line j = value at the flat index of its own pair
    for (int64_t j = 0; j < npairs; ++j) {
      const size_t idx = (size_t)(la[j] - 1) + (size_t)(lb[j] - 1) * NOS;
      szsz[j] = ez[idx];
      spsm[2 * j] = esr[idx];
      spsm[2 * j + 1] = esi[idx];
    }
  });
}

int qs3x_local_mag(qs3x_sector *s, const double *psi, double *expe_sz) {
  return guarded3([&] {
    QS3X_REQUIRE(!is_u1(s), "qs3x_local_mag: not available for U(1)-only sectors");
    QS3X_REQUIRE(s && psi && expe_sz, "qs3x_local_mag: NULL argument");
    qs3x_model *m = s->model;
    const int NOS = m->NOS;
    std::vector<int32_t> off{0}, sites;
    std::vector<std::vector<int>> orbits;
    int no_tot = 0;
    for (int j = 1; j <= NOS; ++j) {
      if (no_tot == NOS) break;
      std::vector<int> lst;
      mk_list_n(m, j, lst);
      if (lst.empty()) continue;
      no_tot += (int)lst.size();
      orbits.push_back(lst);
      for (int v : lst) sites.push_back(v);
      off.push_back((int32_t)sites.size());
    }
    for (int j = 0; j < NOS; ++j) expe_sz[j] = 0.0;
    const int norb = (int)orbits.size();
    if (!norb || !s->dim) return;
    hipStream_t st = stream();
    DevBuf<double> dpsi((size_t)2 * s->dim);
    dpsi.upload(psi, (size_t)2 * s->dim);
    DevBuf<int32_t> doff(off.size()), dsites(sites.size());
    doff.upload(off.data(), off.size());
    dsites.upload(sites.data(), sites.size());
    const int gx = obs_grid(s->dim);
    DevBuf<double> part((size_t)gx * norb);
    hipLaunchKernelGGL(k_lm, dim3(gx, norb), dim3(kObsThreads), 0, st, make_view(s), (const double2 *)dpsi.p, doff.p,
                       dsites.p, part.p);
    QS3X_HIP(hipGetLastError());
    std::vector<double> hp(part.n);
    part.download(hp.data(), hp.size());
    for (int o = 0; o < norb; ++o) {
      double a = 0;
      for (int x = 0; x < gx; ++x) a += hp[(size_t)o * gx + x];
      const int non = (int)orbits[o].size();
      const double val = a / non * 0.5;
      for (int v : orbits[o]) expe_sz[v - 1] = val;
    }
  });
}

int qs3x_apply_sq(qs3x_sector *gs, qs3x_sector *tgt, int32_t spsmsz, const double *q, const double *pos,
                  const double *psi, double *f) {
  return guarded3([&] {
    QS3X_REQUIRE(!is_u1(gs) && !is_u1(tgt), "qs3x_apply_sq: not available for U(1)-only sectors");
    QS3X_REQUIRE(gs && tgt && q && pos && psi && f, "qs3x_apply_sq: NULL argument");
    QS3X_REQUIRE(gs->model == tgt->model, "qs3x_apply_sq: sectors of different models");
    QS3X_REQUIRE(spsmsz >= 1 && spsmsz <= 3, "qs3x_apply_sq: spsmsz must be 1, 2 or 3");
    const int dn = spsmsz == 1 ? -1 : (spsmsz == 2 ? 1 : 0);
    QS3X_REQUIRE(tgt->nod == gs->nod + dn, "qs3x_apply_sq: target N_down does not match spsmsz");
    qs3x_model *m = gs->model;
    const int NOS = m->NOS;
    // exp_list_KS(i) = exp(-i q . R_i) (get_expectation_values.f90:302-304)
    std::vector<double> ek((size_t)2 * NOS);
    for (int i = 0; i < NOS; ++i) {
      const double x = q[0] * pos[3 * i] + q[1] * pos[3 * i + 1] + q[2] * pos[3 * i + 2];
      const std::complex<double> e = std::exp(std::complex<double>(0.0, -x));
      ek[2 * i] = e.real();
      ek[2 * i + 1] = e.imag();
    }
    hipStream_t st = stream();
    DevBuf<double> dek(ek.size()), dpsi((size_t)2 * std::max<int64_t>(gs->dim, 1)),
        df((size_t)2 * std::max<int64_t>(tgt->dim, 1));
    dek.upload(ek.data(), ek.size());
    if (gs->dim) dpsi.upload(psi, (size_t)2 * gs->dim);
    if (tgt->dim) {
      const int gx = obs_grid(tgt->dim);
      hipLaunchKernelGGL(k_sq, dim3(gx), dim3(kObsThreads), 0, st, make_view(gs), make_view(tgt), (int)spsmsz,
                         (const double2 *)dpsi.p, (const double2 *)dek.p, (double2 *)df.p, 1.0 / std::sqrt((double)NOS));
      QS3X_HIP(hipGetLastError());
      df.download(f, (size_t)2 * tgt->dim);
    }
  });
}

// calcu_DSF_wavevector recurrence (get_expectation_values.f90:481-592).  The
// operator applied is A = H - c (c = e_zz for FP32, 0 for FP64); reported
// alpha = <f|A|f> + c - rfield * sztot_new.
int qs3x_dsf_recurrence(qs3x_sector *tgt, int32_t prec, int32_t nvec, const double *f, int32_t itr_dsf,
                        double rfield, double *alpha, double *beta2, double *norm1) {
  return guarded3([&] {
    QS3X_REQUIRE(tgt && f && alpha && beta2 && norm1 && nvec >= 1 && itr_dsf >= 1, "qs3x_dsf_recurrence: bad argument");
    QS3X_REQUIRE(prec == QS3X_FP64 || prec == QS3X_FP32, "qs3x_dsf_recurrence: bad precision");
    const int64_t dim = tgt->dim;
    QS3X_REQUIRE(dim >= 1, "qs3x_dsf_recurrence: empty target sector");
    hipStream_t st = stream();
    const size_t es = prec == QS3X_FP64 ? 16 : 8;
    const double sztot_new = (double)tgt->model->NOS / 2.0 - (double)tgt->nod;
    const double c = (prec == QS3X_FP32) ? tgt->model->ezz : 0.0;
    const int steps = (int)std::min<int64_t>(itr_dsf, dim);
    DevBuf<unsigned char> U[3];
    for (auto &u : U) u.alloc(es * dim);
    DevBuf<double> tmp64((size_t)2 * dim);
    const int NS = steps + 2;
    // slots: arec [NS] | aout [NS] | beta [NS] | scale [NS+1] | norm tmp
    DevBuf<double> sc((size_t)4 * NS + 8);
    const int A_REC = 0, A_OUT = NS, BETA = 2 * NS, SCALE = 3 * NS, TMP = 4 * NS + 2;
    DevBuf<double> part((size_t)2 * kMaxBlocks);
    for (int v = 0; v < nvec; ++v) {
      tmp64.upload(f + (size_t)2 * dim * v, (size_t)2 * dim);
      vec_copy_scale(prec, U[0].p, QS3X_FP64, tmp64.p, dim, nullptr, -1, 1.0, st);
      // norm1 = <f|f> from the FP64 input; scale[0] = 1/sqrt(norm1)
      int nb = vec_dot(QS3X_FP64, tmp64.p, QS3X_FP64, tmp64.p, dim, part.p, st);
      finalize(part.p, nb, sc.p, 0, TMP, -1, 0.0, st);
      finalize(part.p, nb, sc.p, 2, TMP + 1, SCALE, 0.0, st);
      for (int i = 0; i < steps; ++i) {
        void *cur = U[i % 3].p, *prev = U[(i + 2) % 3].p, *w = U[(i + 1) % 3].p;
        launch_hv(tgt, prec, w, cur, sc.p + SCALE + i, part.p, &nb, st);
        finalize(part.p, nb, sc.p, 1, A_REC + i, A_OUT + i, c - rfield * sztot_new, st);
        if (i + 1 == steps) break;
        nb = vec_update3(prec, w, cur, prev, dim, sc.p, A_REC + i, SCALE + i, i > 0 ? BETA + i - 1 : -1,
                         SCALE + i - 1, part.p, st);
        finalize(part.p, nb, sc.p, 2, BETA + i, SCALE + i + 1, 0.0, st);
      }
      std::vector<double> all(sc.n);
      sc.download(all.data(), all.size());
      norm1[v] = all[TMP];
      for (int i = 0; i < itr_dsf; ++i) {
        alpha[(size_t)v * itr_dsf + i] = (i < steps) ? all[A_OUT + i] : 1.0e300;
        beta2[(size_t)v * itr_dsf + i] = (i >= 1 && i < steps) ? all[BETA + i - 1] * all[BETA + i - 1] : 0.0;
      }
    }
  });
}

}  // extern "C"
