// Diagnostics E, I, D, E' as fixed-order FP64 device reductions.
//
// Periodic (flows/monochromatic.py:433-505):
//   E' = 1/2 sum k_metric |u|^2
//   I  = Re( sum_f conj(unit_force_f * F) u[forced_f] ) + I_lam
//   D  = ( sum (-lapl) |u|^2 + 2 I Re - I_lam Re ) / Re   (no k_metric, as
//        the reference sums it)
//   E  = E' - E_lam + I / F   (monochromatic systems; E' otherwise)
// with F = nu (2 pi q_f / Ly)^2 and I_lam = F / 2 (the reference's pi^2/(4 Re)
// and pi^2/(8 Re) at Ly = 4).
// Walls (flows/plane_couette.py:591-613): E' only, with the y quadrature.
#include <hip/hip_runtime.h>

#include <cmath>

#include "context.hpp"
#include "dft.hpp"
#include "passes.hpp"

namespace dnsmi {

struct StatsArgs {
  double F, input_lam, ekin_lam, re;
  int monochromatic;
  int nforced;
  long long idx[4];
  double ufx[4], ufy[4];
};

__global__ void __launch_bounds__(256)
    k_stats_partials(const cx<double>* __restrict__ u, const double* __restrict__ kx_tab,
                     const double* __restrict__ ky_tab, const double* __restrict__ kz_tab, int Nz1,
                     int Nx1, long long n, double* __restrict__ partials) {
  __shared__ double red[8];
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0.0, d = 0.0;
  if (i < n) {
    const int kx = (int)(i % Nx1);
    const long long t = i / Nx1;
    const int kz = (int)(t % Nz1);
    const int s = (int)(t / Nz1);
    const double kxv = kx_tab[kx], kyv = ky_tab[s], kzv = kz_tab[kz];
    const double mlapl = (kxv * kxv + kyv * kyv) + kzv * kzv;
    const double km = kx == 0 ? 1.0 : 2.0;
    for (int c = 0; c < 3; ++c) {
      const cx<double> v = u[c * n + i];
      const double a2 = v.x * v.x + v.y * v.y;
      e += km * a2;
      d += mlapl * a2;
    }
  }
  const double se = block_sum(e, red);
  __syncthreads();
  const double sd = block_sum(d, red + 4);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = se;
    partials[2 * blockIdx.x + 1] = sd;
  }
}

__global__ void __launch_bounds__(64)
    k_stats_finalize(const double* __restrict__ partials, int nb, const cx<double>* __restrict__ u,
                     StatsArgs a, double* __restrict__ out) {
  if (threadIdx.x != 0) return;
  double se = 0.0, sd = 0.0;
  for (int b = 0; b < nb; ++b) {
    se += partials[2 * b];
    sd += partials[2 * b + 1];
  }
  const double ep = se / 2;
  double inp = 0.0, energy = ep;
  if (a.monochromatic) {
    double s = 0.0;
    for (int f = 0; f < a.nforced; ++f) {
      const cx<double> v = u[a.idx[f]];
      const double cfx = a.ufx[f] * a.F, cfy = a.ufy[f] * a.F;  // conj(uf F) v, real part
      s += cfx * v.x + cfy * v.y;
    }
    inp = s + a.input_lam;
    energy = ep - a.ekin_lam + inp / a.F;
  }
  const double enst = sd + 2 * inp * a.re - a.input_lam * a.re;
  out[0] = energy;
  out[1] = inp;
  out[2] = enst / a.re;
  out[3] = ep;
}

int walls_norm(Ctx* c, const double* corr, double* err, hipStream_t s);

__global__ void k_walls_stats_finish(double* out) {
  if (threadIdx.x == 0) {
    const double n = out[3];  // the norm written into slot 3
    out[3] = n * n / 2;
    out[0] = out[3];
    out[1] = 0.0;
    out[2] = 0.0;
  }
}

int get_stats(Ctx* c, const double* state, double* out, hipStream_t s) {
  if (c->walls) {
    // E' = ||u||^2 / 2 with the quadrature norm (the only statistic there)
    int rc = walls_norm(c, state, out + 3, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_walls_stats_finish, dim3(1), dim3(64), 0, s, out);
    return check_launch(c, "walls_stats");
  }
  const long long n = (long long)c->Ny1 * c->Nz1 * c->Nx1;
  const int nb = (int)((n + 255) / 256);
  if (2 * nb > c->max_partials) return c->fail(DNSMI_EINVAL, "stats: too many blocks");
  const dnsmi_config& g = c->cfg;
  StatsArgs a{};
  a.re = g.re;
  a.monochromatic = (g.system == DNSMI_KOLMOGOROV || g.system == DNSMI_WALEFFE);
  if (a.monochromatic) {
    const int qf = 1;
    a.F = (1.0 / g.re) * std::pow(2 * kPi * qf / g.ly, 2);
    a.input_lam = a.F / 2;
    a.ekin_lam = 0.25;
    const double t = kPi * g.tilt_degree / 180;
    const bool tilt = !(std::fabs(g.tilt_degree) == 0 || std::fabs(g.tilt_degree - 180) == 0);
    const bool tilt90 = std::fabs(g.tilt_degree + 90) == 0 || std::fabs(g.tilt_degree - 90) == 0;
    const bool kolm = g.system == DNSMI_KOLMOGOROV;
    // forced modes (flows/monochromatic.py:210-267): component, +-q_f in ky
    int comp[4], qy[4];
    double ux[4], uy[4];
    if (!tilt || tilt90) {
      a.nforced = 2;
      comp[0] = comp[1] = tilt90 ? 2 : 0;
      qy[0] = qf;
      qy[1] = -qf;
      if (kolm) {
        ux[0] = 0; uy[0] = -0.5; ux[1] = 0; uy[1] = 0.5;
      } else {
        ux[0] = 0.5; uy[0] = 0; ux[1] = 0.5; uy[1] = 0;
      }
    } else {
      a.nforced = 4;
      comp[0] = comp[1] = 0;
      comp[2] = comp[3] = 2;
      qy[0] = qy[2] = qf;
      qy[1] = qy[3] = -qf;
      const double ct = std::cos(t), st = std::sin(t);
      if (kolm) {
        ux[0] = 0; uy[0] = -0.5 * ct; ux[1] = 0; uy[1] = 0.5 * ct;
      } else {
        ux[0] = 0.5 * ct; uy[0] = 0; ux[1] = 0.5 * ct; uy[1] = 0;
      }
      ux[2] = 0; uy[2] = -0.5 * st; ux[3] = 0; uy[3] = 0.5 * st;
    }
    for (int f = 0; f < a.nforced; ++f) {
      const int sidx = qy[f] >= 0 ? qy[f] : c->Ny1 + qy[f];
      a.idx[f] = (long long)comp[f] * n + (long long)sidx * c->Nz1 * c->Nx1;
      a.ufx[f] = ux[f];
      a.ufy[f] = uy[f];
    }
  }
  const cx<double>* u = reinterpret_cast<const cx<double>*>(state);
  hipLaunchKernelGGL(k_stats_partials, dim3(nb), dim3(256), 0, s, u, c->kx, c->ky, c->kz, c->Nz1,
                     c->Nx1, n, c->partials);
  int rc = check_launch(c, "stats_partials");
  if (rc) return rc;
  hipLaunchKernelGGL(k_stats_finalize, dim3(1), dim3(64), 0, s, c->partials, nb, u, a, out);
  return check_launch(c, "stats_finalize");
}

}  // namespace dnsmi
