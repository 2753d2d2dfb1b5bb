// Sanitizer driver of the assembly library's host code (the library's host
// pass compiled with -fsanitize=address,undefined; no kernel runs): the
// threaded CSR symbolic phase against a brute-force pattern, and the
// host-side argument checks of lfasm_init / lfasm_init_batch /
// lfasm_residual_static on a handle without a device.  Exit status 0 when
// every check holds; the sanitizers abort on an invalid access.
#include "../../include/lithofem_asm.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <set>
#include <vector>

namespace {

int fails = 0;

void check(bool ok, const char *what)
{
   if (!ok)
   {
      std::printf("FAIL %s\n", what);
      ++fails;
   }
}

}  // namespace

int main()
{
   // ---- lfasm_csr_symbolic: sorted-column union of the element dof blocks
   const int ne = 300, nd = 20, n = 1000;
   std::vector<int> dof((size_t)ne * nd);
   unsigned seed = 12345u;
   for (int e = 0; e < ne; ++e)
   {
      std::set<int> used;
      for (int a = 0; a < nd; ++a)
      {
         int d;
         do
         {
            seed = seed * 1664525u + 1013904223u;
            d = (int)((seed >> 8) % (unsigned)n);
         } while (used.count(d));
         used.insert(d);
         dof[(size_t)e * nd + a] = d;
      }
   }
   std::vector<int> I((size_t)n + 1);
   int *Jc = nullptr;
   const int64_t nnz = lfasm_csr_symbolic(ne, nd, n, dof.data(), I.data(), &Jc);
   std::vector<std::set<int>> rows((size_t)n);
   for (int e = 0; e < ne; ++e)
      for (int a = 0; a < nd; ++a)
         for (int b = 0; b < nd; ++b) { rows[(size_t)dof[(size_t)e * nd + a]].insert(dof[(size_t)e * nd + b]); }
   int64_t want = 0;
   bool same = true;
   for (int r = 0; r < n; ++r)
   {
      if (I[(size_t)r] != want) { same = false; }
      for (int c : rows[(size_t)r])
      {
         if (want >= nnz || Jc[want] != c) { same = false; }
         ++want;
      }
   }
   check(nnz == want && same && I[(size_t)n] == nnz, "csr_symbolic equals the brute-force pattern");
   lfasm_free(Jc);

   // ---- host-side checks of init on a handle without a device
   lfasm_handle *h = nullptr;
   lfasm_create(0, &h);
   check(h != nullptr, "a handle is returned even without a device");
   const int p = 2, nq = 4, nreg = 2;
   std::vector<int> fo((size_t)ne * 4, 1), attr((size_t)ne, 1);
   std::vector<double> sign((size_t)ne * nd, 1.0), Jm((size_t)ne * 9, 0.0), Ji((size_t)ne * 9, 0.0);
   std::vector<double> det((size_t)ne, 1.0), v0((size_t)ne * 3, 0.0);
   std::vector<double> ip((size_t)nq * 3, 0.25), w((size_t)nq, 1.0 / 24), tab((size_t)nq * nd * 3, 0.1);
   std::vector<double> eps(2 * (size_t)nreg, 1.0), TInv(24, 0.0);
   for (int e = 0; e < ne; ++e)
   {
      Jm[(size_t)e * 9] = Jm[(size_t)e * 9 + 4] = Jm[(size_t)e * 9 + 8] = 1.0;
      Ji[(size_t)e * 9] = Ji[(size_t)e * 9 + 4] = Ji[(size_t)e * 9 + 8] = 1.0;
   }
   lfasm_mesh m{};
   m.ne = ne;
   m.nd = nd;
   m.p = p;
   m.height = n;
   m.dof = dof.data();
   m.sign = sign.data();
   m.fo = fo.data();
   m.attr = attr.data();
   m.J = Jm.data();
   m.Jinv = Ji.data();
   m.detJ = det.data();
   m.v0 = v0.data();
   m.rule_cc = {nq, ip.data(), w.data(), tab.data(), tab.data()};
   m.rule_mass = m.rule_cc;
   m.nreg = nreg;
   m.eps_tab = eps.data();
   m.TInv = TInv.data();
   check(lfasm_init(h, &m) == LFASM_E_STATE, "a valid mesh passes validation");
   lfasm_mesh two[2] = {m, m};
   check(lfasm_init_batch(h, two, 2) == LFASM_E_STATE, "a valid batch passes validation");
   std::vector<int> bad = dof;
   bad[5 * nd + 3] = bad[5 * nd + 1];
   m.dof = bad.data();
   check(lfasm_init(h, &m) == LFASM_E_ARG, "a repeated dof is refused");
   m.dof = dof.data();
   attr[7] = nreg + 1;
   check(lfasm_init(h, &m) == LFASM_E_ARG, "an attribute outside the table is refused");
   attr[7] = 1;
   fo[9] = 6;
   check(lfasm_init(h, &m) == LFASM_E_ARG, "a face orientation outside 0..5 is refused");
   fo[9] = 1;
   det[3] = -1.0;
   check(lfasm_init(h, &m) == LFASM_E_ARG, "an inverted element is refused");
   det[3] = 1.0;
   m.nd = 21;
   check(lfasm_init(h, &m) == LFASM_E_UNSUPPORTED, "nd inconsistent with p is refused");
   m.nd = nd;
   check(lfasm_init_vertices(h, &m, nullptr) == LFASM_E_ARG, "init_vertices needs coordinates");
   std::vector<double> x(8, 0.0);
   double rel = 0.0;
   check(lfasm_residual_static(h, 2 * (int64_t)(LFASM_RESIDUAL_STATIC_MAX_N + 1), x.data(), x.data(),
                               &rel) == LFASM_E_ARG, "residual_static refuses sizes beyond its workspace");
   check(lfasm_residual_static(h, 7, x.data(), x.data(), &rel) == LFASM_E_ARG, "odd vector length refused");
   check(lfasm_pml_smax(1e-8, 2.0, 0.4654211, 1.0, 13.5) > 4.3, "pml smax");
   lfasm_destroy(h);
   std::printf("san_lib %s (%d failed checks)\n", fails ? "FAILED" : "ok", fails);
   return fails ? 1 : 0;
}
