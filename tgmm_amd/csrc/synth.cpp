// synth.cpp -- seeded synthetic microscopy frames in the reference's
// boundary layout (host C++, OpenMP). Used by tests and the benchmark to
// build inputs of the shapes the EM is measured on; not part of the EM.
//
// A frame is a set of cells, each an anisotropic 3D Gaussian blob (physical
// radii r1..r3, a uniformly random rotation, z sampled every `aniso_z`
// physical units). Voxels within Mahalanobis radius rho of a centre are
// foreground; a voxel claimed by several cells goes to the one with the
// smallest Mahalanobis distance. Intensity is
//   u16 = Poisson(1000 * brightness * exp(-d^2/2) + 100).
// Each cell's voxels are split into supervoxels by recursive median cuts
// along the longest physical axis until at most `max_sv_size` remain.
// Preprocessing follows the tracker (main.cpp:598-669, 774-870):
// per-supervoxel minimum subtracted, frame scaled to [0,1] as float `img`,
// integer voxel coordinates as float SoA `query`, voxels contiguous per
// supervoxel with `sv_ptr`, and an intensity-weighted centroid per
// supervoxel (supervoxel.cpp:1410-1449; unweighted mean when the weight sum
// is below 1e-5).
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <parallel/algorithm>
#include <vector>

#include "../../include/tgmm_em.h"

namespace {

inline uint64_t splitmix64(uint64_t& x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  double uni() { return (splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }
  double normal() {
    double u1 = uni(), u2 = uni();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
  uint32_t poisson(double lam) {
    if (lam <= 0) return 0;
    if (lam < 30.0) {
      double L = std::exp(-lam), p = 1.0;
      uint32_t k = 0;
      do {
        k++;
        p *= uni();
      } while (p > L);
      return k - 1;
    }
    double v = lam + std::sqrt(lam) * normal() + 0.5;
    return v < 0 ? 0u : (uint32_t)v;
  }
};

inline uint64_t mix(uint64_t a, uint64_t b) {
  uint64_t x = a ^ (b * 0x9E3779B97F4A7C15ull);
  return splitmix64(x);
}

struct Cell {
  double c[3];       // centre, voxel units
  double P[6];       // precision in physical units (packed)
  double bright;
  double ext[3];     // half extent in voxel units at rho
};

struct Claim {
  uint64_t key;
  float d2;
  int cell;
};

}  // namespace

extern "C" {

typedef struct tgmm_synth_cells {
  int n;
  double* centre;  // [3n] voxel units
  double* radius;  // [3n] physical units
  double* quat;    // [4n] unit quaternion
  double* bright;  // [n]
} tgmm_synth_cells;

typedef struct tgmm_synth_frame {
  long long n_voxels, n_supervoxels;
  int n_cells;
  float* query;       // [3N] SoA
  float* img;         // [N]
  float* centroids;   // [3S] SoA
  long long* sv_ptr;  // [S+1]
  int* sv_cell;       // [S] owning cell
  unsigned short* raw;  // [N] raw uint16 intensities
} tgmm_synth_frame;

// Sample `n` cells uniformly in vol (voxel units); radii U[rmin, rmax]
// (physical), uniform random rotations, brightness lognormal(0, bsig) (1 when
// bsig == 0). A fraction `pair_frac` of cells is placed as close pairs whose
// centres are U[pd_min, pd_max] voxels apart.
int tgmm_synth_sample_cells(uint64_t seed, int n, const double* vol,
                            double aniso_z, double rmin, double rmax,
                            double pair_frac, double pd_min, double pd_max,
                            double bsig, tgmm_synth_cells* out) {
  if (n < 0 || !out) return TGMM_ERR_ARG;
  out->n = n;
  Rng r(seed);
  int npairs = (int)std::floor(n * pair_frac * 0.5);
  for (int k = 0; k < n; k++) {
    double* c = out->centre + 3 * k;
    if (k < 2 * npairs && (k & 1)) {
      // partner of cell k-1: random direction in physical space
      double u[3], nn = 0;
      for (int i = 0; i < 3; i++) {
        u[i] = r.normal();
        nn += u[i] * u[i];
      }
      nn = std::sqrt(nn);
      double dist = pd_min + (pd_max - pd_min) * r.uni();
      const double* p = out->centre + 3 * (k - 1);
      c[0] = p[0] + dist * u[0] / nn;
      c[1] = p[1] + dist * u[1] / nn;
      c[2] = p[2] + dist * u[2] / nn / aniso_z;
      for (int i = 0; i < 3; i++) c[i] = std::min(std::max(c[i], 0.0), vol[i] - 1.0);
    } else {
      for (int i = 0; i < 3; i++) c[i] = r.uni() * (vol[i] - 1.0);
    }
    for (int i = 0; i < 3; i++) out->radius[3 * k + i] = rmin + (rmax - rmin) * r.uni();
    double q[4], qn = 0;
    for (int i = 0; i < 4; i++) {
      q[i] = r.normal();
      qn += q[i] * q[i];
    }
    qn = std::sqrt(qn);
    for (int i = 0; i < 4; i++) out->quat[4 * k + i] = q[i] / qn;
    out->bright[k] = bsig > 0 ? std::exp(bsig * r.normal()) : 1.0;
  }
  return TGMM_OK;
}

// Rasterise cells into a frame. target_voxels sets the Mahalanobis radius
// rho so that the expected foreground count is near the target (overlaps
// lower it slightly). Arrays are allocated with new[]; free with
// tgmm_synth_free.
int tgmm_synth_rasterize(uint64_t seed, const tgmm_synth_cells* cells,
                         const double* vol, double aniso_z,
                         double target_voxels, int max_sv_size,
                         tgmm_synth_frame* out) {
  if (!cells || !out || cells->n < 0 || max_sv_size < 1) return TGMM_ERR_ARG;
  const int n = cells->n;
  std::vector<Cell> cs(n);
  double vsum = 0.0;
  for (int k = 0; k < n; k++) {
    const double* rr = cells->radius + 3 * k;
    vsum += 4.18879020478639 * rr[0] * rr[1] * rr[2] / aniso_z;
  }
  const double rho = vsum > 0 ? std::cbrt(std::max(target_voxels, 1.0) / vsum) : 1.0;
  for (int k = 0; k < n; k++) {
    const double* q = cells->quat + 4 * k;
    const double w = q[0], x = q[1], y = q[2], z = q[3];
    const double R[3][3] = {{1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)},
                            {2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)},
                            {2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)}};
    double Cov[3][3], Prec[3][3];
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) {
        double sc = 0, sp = 0;
        for (int i = 0; i < 3; i++) {
          const double ri = cells->radius[3 * k + i];
          sc += R[a][i] * R[b][i] * ri * ri;
          sp += R[a][i] * R[b][i] / (ri * ri);
        }
        Cov[a][b] = sc;
        Prec[a][b] = sp;
      }
    Cell& c = cs[k];
    for (int i = 0; i < 3; i++) c.c[i] = cells->centre[3 * k + i];
    int m = 0;
    for (int a = 0; a < 3; a++)
      for (int b = a; b < 3; b++) c.P[m++] = Prec[a][b];
    c.bright = cells->bright[k];
    c.ext[0] = rho * std::sqrt(Cov[0][0]);
    c.ext[1] = rho * std::sqrt(Cov[1][1]);
    c.ext[2] = rho * std::sqrt(Cov[2][2]) / aniso_z;
  }
  const int64_t X = (int64_t)vol[0], Y = (int64_t)vol[1], Z = (int64_t)vol[2];
  const double rho2 = rho * rho;
  // per-cell candidate claims
  std::vector<std::vector<Claim>> per(n);
#pragma omp parallel for schedule(dynamic, 32)
  for (int k = 0; k < n; k++) {
    const Cell& c = cs[k];
    int64_t lo[3], hi[3];
    const int64_t dims[3] = {X, Y, Z};
    for (int i = 0; i < 3; i++) {
      lo[i] = std::max<int64_t>(0, (int64_t)std::floor(c.c[i] - c.ext[i]));
      hi[i] = std::min<int64_t>(dims[i] - 1, (int64_t)std::ceil(c.c[i] + c.ext[i]));
    }
    std::vector<Claim>& v = per[k];
    for (int64_t zz = lo[2]; zz <= hi[2]; zz++)
      for (int64_t yy = lo[1]; yy <= hi[1]; yy++)
        for (int64_t xx = lo[0]; xx <= hi[0]; xx++) {
          const double d0 = xx - c.c[0], d1 = yy - c.c[1], d2 = (zz - c.c[2]) * aniso_z;
          const double q = c.P[0] * d0 * d0 + c.P[3] * d1 * d1 + c.P[5] * d2 * d2 +
                           2 * (c.P[1] * d0 * d1 + c.P[2] * d0 * d2 + c.P[4] * d1 * d2);
          if (q <= rho2) v.push_back({(uint64_t)((zz * Y + yy) * X + xx), (float)q, k});
        }
  }
  std::vector<size_t> off(n + 1, 0);
  for (int k = 0; k < n; k++) off[k + 1] = off[k] + per[k].size();
  std::vector<Claim> all(off[n]);
#pragma omp parallel for
  for (int k = 0; k < n; k++) std::copy(per[k].begin(), per[k].end(), all.begin() + off[k]);
  per.clear();
  per.shrink_to_fit();
  __gnu_parallel::sort(all.begin(), all.end(), [](const Claim& a, const Claim& b) {
    if (a.key != b.key) return a.key < b.key;
    if (a.d2 != b.d2) return a.d2 < b.d2;
    return a.cell < b.cell;
  });
  // keep the best claim per voxel, then regroup per cell in key order
  std::vector<int> cnt(n + 1, 0);
  std::vector<Claim> kept;
  kept.reserve(all.size());
  for (size_t i = 0; i < all.size(); i++)
    if (i == 0 || all[i].key != all[i - 1].key) {
      kept.push_back(all[i]);
      cnt[all[i].cell + 1]++;
    }
  all.clear();
  all.shrink_to_fit();
  for (int k = 0; k < n; k++) cnt[k + 1] += cnt[k];
  std::vector<Claim> bycell(kept.size());
  {
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (const Claim& c : kept) bycell[fill[c.cell]++] = c;
  }
  kept.clear();
  kept.shrink_to_fit();
  const int64_t N = (int64_t)bycell.size();
  // supervoxels per cell: recursive median split (in place on bycell ranges)
  std::vector<std::vector<std::pair<int64_t, int64_t>>> svs(n);
  auto coord = [&](uint64_t key, int ax) -> double {
    int64_t x = key % X, y = (key / X) % Y, z = key / (X * Y);
    return ax == 0 ? (double)x : ax == 1 ? (double)y : (double)z * aniso_z;
  };
#pragma omp parallel for schedule(dynamic, 32)
  for (int k = 0; k < n; k++) {
    std::vector<std::pair<int64_t, int64_t>> stack{{cnt[k], cnt[k + 1]}};
    std::vector<std::pair<int64_t, int64_t>>& res = svs[k];
    while (!stack.empty()) {
      auto rg = stack.back();
      stack.pop_back();
      const int64_t len = rg.second - rg.first;
      if (len <= 0) continue;
      if (len <= max_sv_size) {
        res.push_back(rg);
        continue;
      }
      double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
      for (int64_t i = rg.first; i < rg.second; i++)
        for (int a = 0; a < 3; a++) {
          double v = coord(bycell[i].key, a);
          lo[a] = std::min(lo[a], v);
          hi[a] = std::max(hi[a], v);
        }
      int ax = 0;
      for (int a = 1; a < 3; a++)
        if (hi[a] - lo[a] > hi[ax] - lo[ax]) ax = a;
      std::stable_sort(bycell.begin() + rg.first, bycell.begin() + rg.second,
                       [&](const Claim& a, const Claim& b) {
                         double va = coord(a.key, ax), vb = coord(b.key, ax);
                         if (va != vb) return va < vb;
                         return a.key < b.key;
                       });
      const int64_t mid = rg.first + len / 2;
      stack.push_back({mid, rg.second});  // right processed after left
      stack.push_back({rg.first, mid});
    }
    for (auto& rg : res)
      std::sort(bycell.begin() + rg.first, bycell.begin() + rg.second,
                [](const Claim& a, const Claim& b) { return a.key < b.key; });
  }
  int64_t S = 0;
  for (int k = 0; k < n; k++) S += (int64_t)svs[k].size();
  out->n_voxels = N;
  out->n_supervoxels = S;
  out->n_cells = n;
  out->query = new float[3 * std::max<int64_t>(N, 1)];
  out->img = new float[std::max<int64_t>(N, 1)];
  out->raw = new unsigned short[std::max<int64_t>(N, 1)];
  out->centroids = new float[3 * std::max<int64_t>(S, 1)];
  out->sv_ptr = new long long[S + 1];
  out->sv_cell = new int[std::max<int64_t>(S, 1)];
  // ordering: cells in index order, supervoxels in split order
  std::vector<int64_t> svbase(n + 1, 0);
  for (int k = 0; k < n; k++) svbase[k + 1] = svbase[k] + (int64_t)svs[k].size();
  std::vector<int64_t> svlen(S);
  for (int k = 0; k < n; k++)
    for (size_t i = 0; i < svs[k].size(); i++) {
      svlen[svbase[k] + i] = svs[k][i].second - svs[k][i].first;
      out->sv_cell[svbase[k] + i] = k;
    }
  out->sv_ptr[0] = 0;
  for (int64_t s = 0; s < S; s++) out->sv_ptr[s + 1] = out->sv_ptr[s] + svlen[s];
  std::vector<uint32_t> wsub(N);
  uint32_t wmax = 0;
#pragma omp parallel for schedule(dynamic, 32) reduction(max : wmax)
  for (int k = 0; k < n; k++) {
    for (size_t i = 0; i < svs[k].size(); i++) {
      const int64_t s = svbase[k] + (int64_t)i;
      int64_t v = out->sv_ptr[s];
      uint32_t mn = 0xFFFFFFFFu;
      const int64_t b0 = svs[k][i].first, b1 = svs[k][i].second;
      for (int64_t j = b0; j < b1; j++, v++) {
        const Claim& c = bycell[j];
        const uint64_t key = c.key;
        out->query[v] = (float)(key % X);
        out->query[v + N] = (float)((key / X) % Y);
        out->query[v + 2 * N] = (float)(key / (X * Y));
        Rng rr(mix(seed, key));
        const double lam = 1000.0 * cs[k].bright * std::exp(-0.5 * (double)c.d2) + 100.0;
        uint32_t u = std::min<uint32_t>(rr.poisson(lam), 65535u);
        out->raw[v] = (unsigned short)u;
        mn = std::min(mn, u);
      }
      double sw = 0, sx[3] = {0, 0, 0}, ux[3] = {0, 0, 0};
      for (int64_t vv = out->sv_ptr[s]; vv < out->sv_ptr[s + 1]; vv++) {
        const uint32_t w = out->raw[vv] - mn;
        wsub[vv] = w;
        wmax = std::max(wmax, w);
        for (int a = 0; a < 3; a++) {
          sx[a] += (double)w * out->query[vv + a * N];
          ux[a] += out->query[vv + a * N];
        }
        sw += w;
      }
      const double len = (double)(out->sv_ptr[s + 1] - out->sv_ptr[s]);
      for (int a = 0; a < 3; a++)
        out->centroids[s + a * S] = (float)(sw >= 1e-5 ? sx[a] / sw : ux[a] / len);
    }
  }
  const double scale = wmax > 0 ? 1.0 / (double)wmax : 0.0;
#pragma omp parallel for
  for (int64_t v = 0; v < N; v++) out->img[v] = (float)(wsub[v] * scale);
  return TGMM_OK;
}

// Initial Gaussians, one per cell: m_o = m_k = centre + N(0, sd^2) per axis,
// priors alpha_o, beta_o, nu_o, W_o = w_diag * I; posterior = prior.
int tgmm_synth_init_gaussians(uint64_t seed, const tgmm_synth_cells* cells,
                              double sd, double alpha_o, double beta_o,
                              double nu_o, double w_diag,
                              GaussianMixtureModelCUDA* gm) {
  if (!cells || !gm) return TGMM_ERR_ARG;
  Rng r(seed);
  for (int k = 0; k < cells->n; k++) {
    GaussianMixtureModelCUDA& g = gm[k];
    std::memset(&g, 0, sizeof(g));
    for (int i = 0; i < 3; i++) {
      g.m_o[i] = cells->centre[3 * k + i] + sd * r.normal();
      g.m_k[i] = g.m_o[i];
    }
    g.alpha_o = g.alpha_k = alpha_o;
    g.beta_o = g.beta_k = beta_o;
    g.nu_o = g.nu_k = nu_o;
    const double W[6] = {w_diag, 0, 0, w_diag, 0, w_diag};
    for (int i = 0; i < 6; i++) g.W_o[i] = g.W_k[i] = W[i];
  }
  return TGMM_OK;
}

void tgmm_synth_free(tgmm_synth_frame* f) {
  if (!f) return;
  delete[] f->query;
  delete[] f->img;
  delete[] f->raw;
  delete[] f->centroids;
  delete[] f->sv_ptr;
  delete[] f->sv_cell;
  std::memset(f, 0, sizeof(*f));
}

int tgmm_synth_record_size(void) { return (int)sizeof(GaussianMixtureModelCUDA); }
}
