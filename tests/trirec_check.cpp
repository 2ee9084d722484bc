// Checks the invariants of the per-triangle bounding sphere that the closest-point pruning bound
// relies on (thea_amd/csrc/tri_math.hpp make_trirec), on seeded random triangles: ordinary ones,
// slivers, collinear-in-reals triples perturbed by float rounding, repeated vertices, near and far
// from the origin. Every vertex lies within the radius |r| of the centre (double arithmetic);
// where the record marks its plane tight (r >= 0), |n|^2 is within 2^-18 of 1, every vertex is
// within 2^-20 of its largest coordinate of the plane n.x = pd, and so is the centre; where it does
// not (r < 0), the centre is the box centre. Prints "<triangles> <violations> <tight> <not tight>".
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "tri_math.hpp"

int main(int argc, char **argv) {
  const long trials = argc > 1 ? std::atol(argv[1]) : 200000;
  std::mt19937_64 g(7);
  std::normal_distribution<double> N(0.0, 1.0);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  long bad = 0, tight = 0, loose = 0;
  for (long t = 0; t < trials; ++t) {
    const double scale_off[4] = {0.0, 1.0, 1e3, 1e5};
    const double off = scale_off[t % 4];
    const double size = std::pow(10.0, -4.0 + 5.0 * U(g));
    double o[3], d[3], e[3];
    for (int a = 0; a < 3; ++a) { o[a] = off * N(g); d[a] = size * N(g); e[a] = size * N(g); }
    float3 v0 = thg::f3((float)o[0], (float)o[1], (float)o[2]), v1, v2;
    switch ((t / 4) % 4) {
      case 0:  // ordinary
        v1 = thg::f3((float)(o[0] + d[0]), (float)(o[1] + d[1]), (float)(o[2] + d[2]));
        v2 = thg::f3((float)(o[0] + e[0]), (float)(o[1] + e[1]), (float)(o[2] + e[2]));
        break;
      case 1:  // sliver
        v1 = thg::f3((float)(o[0] + d[0]), (float)(o[1] + d[1]), (float)(o[2] + d[2]));
        v2 = thg::f3((float)(o[0] + 0.5 * d[0] + 1e-4 * e[0]), (float)(o[1] + 0.5 * d[1] + 1e-4 * e[1]),
                     (float)(o[2] + 0.5 * d[2] + 1e-4 * e[2]));
        break;
      case 2:  // collinear in real numbers
        v1 = thg::f3((float)(o[0] + d[0]), (float)(o[1] + d[1]), (float)(o[2] + d[2]));
        v2 = thg::f3((float)(o[0] + 2 * d[0]), (float)(o[1] + 2 * d[1]), (float)(o[2] + 2 * d[2]));
        break;
      default:  // repeated vertex
        v1 = v0;
        v2 = thg::f3((float)(o[0] + e[0]), (float)(o[1] + e[1]), (float)(o[2] + e[2]));
        break;
    }
    const thg::TriRec r = thg::make_trirec(v0, v1, v2, (int32_t)t);
    const double c[3] = {r.sb.x, r.sb.y, r.sb.z};
    const double rad = std::fabs((double)r.sb.w);
    const float3 vv[3] = {v0, v1, v2};
    bool ok = true;
    double vmax = 0;
    for (int k = 0; k < 3; ++k) {
      const double dx = vv[k].x - c[0], dy = vv[k].y - c[1], dz = vv[k].z - c[2];
      ok = ok && std::sqrt(dx * dx + dy * dy + dz * dz) <= rad;
      vmax = std::fmax(vmax, std::fmax(std::fabs(vv[k].x), std::fmax(std::fabs(vv[k].y), std::fabs(vv[k].z))));
    }
    const double n[3] = {r.p0.w, r.p1.w, r.p2.w}, pd = r.pl.x;
    if (!std::signbit(r.sb.w)) {
      ++tight;
      const double nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
      ok = ok && std::fabs(nn - 1.0) <= 1.0 / 262144.0;
      for (int k = 0; k < 3; ++k)
        ok = ok && std::fabs(n[0] * vv[k].x + n[1] * vv[k].y + n[2] * vv[k].z - pd) <= vmax / 1048576.0;
      ok = ok && std::fabs(n[0] * c[0] + n[1] * c[1] + n[2] * c[2] - pd) <= vmax / 1048576.0 + 1e-30;
    } else {
      ++loose;
      for (int a = 0; a < 3; ++a) {
        const double lo = std::fmin((&v0.x)[a], std::fmin((&v1.x)[a], (&v2.x)[a]));
        const double hi = std::fmax((&v0.x)[a], std::fmax((&v1.x)[a], (&v2.x)[a]));
        ok = ok && c[a] == (double)(float)(0.5 * (lo + hi));
      }
    }
    if (!ok) {
      if (bad < 5) std::fprintf(stderr, "violation at trial %ld\n", t);
      ++bad;
    }
  }
  std::printf("%ld %ld %ld %ld\n", trials, bad, tight, loose);
  return 0;
}
