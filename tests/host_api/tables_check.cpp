// Host-side table and geometry builders of the library (hd_tables.cpp,
// hd_geometry.cpp), exercised over every supported degree, point count and
// dimension; built with AddressSanitizer and UndefinedBehaviorSanitizer by
// tests/test_sanitizers.py.  Prints a checksum line and checks invariants
// that hold for every table: Sinv S = I, the collapsed 1D operator
// annihilates constants, and |det J| w sums to the cell volume.
#include "hd_geometry.h"
#include "hd_tables.h"

#include <cmath>
#include <cstdio>
#include <vector>

int main()
{
  double sum = 0.0;
  int bad = 0;
  for (int k = 1; k <= 7; ++k)
    for (int nq = k + 1; nq <= 8; ++nq)
      for (int coll = 0; coll < 2; ++coll)
        {
          if (coll && nq != k + 1)
            continue;
          hd::Tables1D t;
          if (!t.build(k, nq, coll != 0))
            continue;
          const int nd = k + 1;
          // Sinv S = I (nd x nd)
          for (int i = 0; i < nd; ++i)
            for (int j = 0; j < nd; ++j)
              {
                double s = 0.0;
                for (int q = 0; q < nq; ++q)
                  s += t.Sinv[i * nq + q] * t.S[q * nd + j];
                if (std::fabs(s - (i == j ? 1.0 : 0.0)) > 1e-10)
                  ++bad;
              }
          for (double a : {0.7, -0.4, 0.0})
            {
              const hd::Collapsed1D c = hd::collapse_direction(t, a, 0.25, 0.5);
              // constants are steady: A 1 + b = 0
              for (int i = 0; i < nd; ++i)
                {
                  double s = c.b[i];
                  for (int j = 0; j < nd; ++j)
                    s += c.A[i * nd + j];
                  if (std::fabs(s) > 1e-9 * (1.0 + std::fabs(a)) * 64)
                    ++bad;
                  sum += s;
                }
            }
          sum += t.S[0] + t.Dq[0];
        }
  for (int dim = 1; dim <= 3; ++dim)
    {
      hd::SpaceMap m;
      m.dim = dim;
      for (int d = 0; d < dim; ++d)
        {
          m.n[d] = 3;
          m.lo[d] = -1.0;
          m.hi[d] = 1.0;
          m.h[d] = 2.0 / 3;
        }
      m.delta = 0.1;
      m.freq = 2;
      hd::Tables1D t;
      t.build(3, 4, false);
      const double a[3] = {0.3, -0.2, 0.5};
      std::vector<double> cell, face;
      hd::space_tables(m, 4, t.qpoints.data(), t.qweights.data(), a, cell,
                       face);
      double vol = 0.0;
      for (size_t i = dim; i < cell.size(); i += dim + 1)
        vol += cell[i];
      if (std::fabs(vol - std::pow(2.0, dim)) > 1e-6)
        ++bad;
      for (double f : face)
        sum += f;
      sum += hd::space_max_speed(m, 4, t.qpoints.data(), a);
    }
  std::printf("checksum %.12e bad %d\n", sum, bad);
  return bad == 0 ? 0 : 1;
}
