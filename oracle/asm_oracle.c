/*
 * CPU oracle for the Nedelec system-matrix assembly path.
 *
 * TEST INFRASTRUCTURE ONLY (tests/, smoke(), bench.py cpu_baseline).  A
 * plain-C restatement of the reference's CPU reassembler, which the
 * reference itself pins to MFEM's SesquilinearForm::Assemble(0) at 1e-13
 * (tests/test_asm_v25.py:115-117).  Operation order follows the cited
 * reference lines term by term; complex arithmetic uses C99 double _Complex,
 * whose real*complex scaling and complex*complex / real/complex operations
 * are the same libgcc operations std::complex<double> compiles to.
 *
 * Build: oracle/Makefile (-O2 -ffp-contract=off, so results do not depend
 * on the host's FMA support).
 */
#include "asm_oracle.h"

#include <complex.h>
#undef I
#define IU _Complex_I
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double _Complex C;

static C ld(const double *p) { return p[0] + IU * p[1]; }

/* TInv table of solver/asm_data.inc:26-34 (column-major 2x2 per
 * orientation).  Pinned by tests against T*TInv = I and against the face
 * DOF conformity of the standalone basis. */
static const double orc_TInv[24] = {
   1.0, 0.0, 0.0, 1.0,
   -1.0, -1.0, 0.0, 1.0,
   -1.0, -1.0, 1.0, 0.0,
   1.0, 0.0, -1.0, -1.0,
   0.0, 1.0, -1.0, -1.0,
   0.0, 1.0, 1.0, 0.0};

/* solver/problem_common.inc:69-90: PML stretch with n_med found by
 * scanning the region table for the side's PML region. */
static C s_z(const orc_problem *pb, double z)
{
   double d = -1.0, t = pb->pml_thick;
   double n_med = 1.0;
   if (z < pb->z_min)
   {
      d = (pb->z_min - z) / t;
      for (int r = 0; r < pb->nreg; ++r)
         if (pb->is_pml[r] && !pb->pml_top[r]) { n_med = cabs(csqrt(ld(pb->eps + 2 * r))); }
   }
   else if (z > pb->z_max)
   {
      d = (z - pb->z_max) / t;
      for (int r = 0; r < pb->nreg; ++r)
         if (pb->is_pml[r] && pb->pml_top[r]) { n_med = cabs(csqrt(ld(pb->eps + 2 * r))); }
   }
   if (d < 0.0) { return 1.0 + IU * 0.0; }
   if (d > 1.0) { d = 1.0; }
   const double smax = -log(pb->pml_target) * (pb->pml_order + 1) /
                       (2.0 * pb->k0 * n_med * pb->pml_thick);
   return 1.0 + IU * (smax * pow(d, pb->pml_order));
}

void orc_s_z(const orc_problem *pb, double z, double out[2])
{
   const C s = s_z(pb, z);
   out[0] = creal(s);
   out[1] = cimag(s);
}

/* solver/problem_common.inc:92-98 */
static void lam_inv_diag(const orc_problem *pb, double z, C m[3])
{
   const C s = s_z(pb, z);
   m[0] = 1.0 / s;
   m[1] = 1.0 / s;
   m[2] = s;
}

static void reassemble_element_eps(const orc_asm_data *ad, const orc_problem *pb,
                                   int e, int mask, int apply_te, const C *eps_override,
                                   C *Ae, double *pc, double *pv)
{
   const int nd = ad->nd;
   const double *J = ad->J + (size_t)e * 9;
   const double *Ji = ad->Jinv + (size_t)e * 9;
   const double detJ = ad->detJ[e];
   const double *v0 = ad->v0 + (size_t)e * 3;
   C eps;
   if (eps_override) { eps = *eps_override; }
   else if (ad->eps_elem) { eps = ld(ad->eps_elem + 2 * (size_t)e); }
   else { eps = ld(pb->eps + 2 * (ad->attr[e] - 1)); }
   const double k2 = pb->k0 * pb->k0;

   for (size_t i = 0; i < (size_t)nd * nd; ++i) { Ae[i] = 0.0; }

   /* curl-curl term on the 2p-2 rule (solver/asm_data.inc:231-259) */
   if (mask & ORC_CURLCURL)
   {
      const orc_rule *rc = &ad->rule_cc;
      for (int q = 0; q < rc->nq; ++q)
      {
         const double *xi = rc->ip + (size_t)q * 3;
         const double z = v0[2] + J[6] * xi[0] + J[7] * xi[1] + J[8] * xi[2];
         C li[3];
         lam_inv_diag(pb, z, li);
         const double w = rc->w[q] * detJ;
         const double *cs = rc->curl + (size_t)q * nd * 3;
         for (int k = 0; k < nd; ++k)
            for (int i = 0; i < 3; ++i)
            {
               pc[(size_t)k * 3 + i] =
                  (cs[k * 3 + 0] * J[i * 3 + 0] + cs[k * 3 + 1] * J[i * 3 + 1] +
                   cs[k * 3 + 2] * J[i * 3 + 2]) / detJ;
            }
         for (int k = 0; k < nd; ++k)
            for (int l = 0; l < nd; ++l)
            {
               C s = 0.0;
               for (int i = 0; i < 3; ++i)
               {
                  s += pc[(size_t)k * 3 + i] * li[i] * pc[(size_t)l * 3 + i];
               }
               Ae[(size_t)k * nd + l] += w * s;
            }
      }
   }

   /* mass + k-parallel terms on the 2p rule (solver/asm_data.inc:261-335) */
   const int mmask = mask & (ORC_MASS | ORC_CROSS0 | ORC_CROSS1 | ORC_KLK);
   if (mmask)
   {
      const orc_rule *rm = &ad->rule_mass;
      double K[9];
      K[0] = 0.0;      K[1] = 0.0;      K[2] = pb->ky;
      K[3] = 0.0;      K[4] = 0.0;      K[5] = -pb->kx;
      K[6] = -pb->ky;  K[7] = pb->kx;   K[8] = 0.0;
      for (int q = 0; q < rm->nq; ++q)
      {
         const double *xi = rm->ip + (size_t)q * 3;
         const double z = v0[2] + J[6] * xi[0] + J[7] * xi[1] + J[8] * xi[2];
         const C s = s_z(pb, z);
         const C lam[3] = {s, s, 1.0 / s};
         C li[3];
         lam_inv_diag(pb, z, li);
         const double w = rm->w[q] * detJ;
         const double *vs = rm->vshape + (size_t)q * nd * 3;
         const double *cs = rm->curl + (size_t)q * nd * 3;
         for (int k = 0; k < nd; ++k)
            for (int i = 0; i < 3; ++i)
            {
               pv[(size_t)k * 3 + i] =
                  vs[k * 3 + 0] * Ji[0 * 3 + i] + vs[k * 3 + 1] * Ji[1 * 3 + i] +
                  vs[k * 3 + 2] * Ji[2 * 3 + i];
               if (ad->has_kpar)
               {
                  pc[(size_t)k * 3 + i] =
                     (cs[k * 3 + 0] * J[i * 3 + 0] + cs[k * 3 + 1] * J[i * 3 + 1] +
                      cs[k * 3 + 2] * J[i * 3 + 2]) / detJ;
               }
            }
         if (mask & ORC_MASS)
         {
            const C mdiag[3] = {-k2 * eps * lam[0], -k2 * eps * lam[1],
                                -k2 * eps * lam[2]};
            for (int k = 0; k < nd; ++k)
               for (int l = 0; l < nd; ++l)
               {
                  C s2 = 0.0;
                  for (int i = 0; i < 3; ++i)
                  {
                     s2 += pv[(size_t)k * 3 + i] * mdiag[i] * pv[(size_t)l * 3 + i];
                  }
                  Ae[(size_t)k * nd + l] += w * s2;
               }
         }
         if (ad->has_kpar && (mask & (ORC_CROSS0 | ORC_CROSS1 | ORC_KLK)))
         {
            const C iu = IU;
            for (int k = 0; k < nd; ++k)
               for (int l = 0; l < nd; ++l)
               {
                  C s0 = 0.0, s1 = 0.0, s2b = 0.0;
                  for (int r = 0; r < 3; ++r)
                     for (int c = 0; c < 3; ++c)
                     {
                        const double krc = K[r * 3 + c];
                        if (krc != 0.0)
                        {
                           if (mask & ORC_CROSS0)
                           {
                              s0 += pv[(size_t)k * 3 + r] * (iu * krc * li[c]) *
                                    pc[(size_t)l * 3 + c];
                           }
                           if (mask & ORC_CROSS1)
                           {
                              s1 += pc[(size_t)k * 3 + r] * (iu * li[r] * krc) *
                                    pv[(size_t)l * 3 + c];
                           }
                        }
                        if (mask & ORC_KLK)
                        {
                           C m2 = 0.0;
                           for (int t = 0; t < 3; ++t)
                           {
                              m2 += K[r * 3 + t] * li[t] * K[t * 3 + c];
                           }
                           s2b -= pv[(size_t)k * 3 + r] * m2 * pv[(size_t)l * 3 + c];
                        }
                     }
                  Ae[(size_t)k * nd + l] += w * (s0 + s1 + s2b);
               }
         }
      }
   }

   /* dual DofTransformation TInv(fo)^T on face-DOF pairs, columns then rows
    * (solver/asm_data.inc:337-369) */
   if (apply_te && ad->p >= 2)
   {
      const int p = ad->p;
      const int foff = 6 * p;
      const int ntd = p * (p - 1);
      const int *fo = ad->fo + (size_t)e * 4;
      for (int axis = 0; axis < 2; ++axis)
      {
         for (int f = 0; f < 4; ++f)
         {
            const double *Ti = orc_TInv + 4 * fo[f];
            for (int i = 0; i < ntd / 2; ++i)
            {
               const int d0 = foff + f * ntd + 2 * i;
               for (int r = 0; r < nd; ++r)
               {
                  C *a = axis == 0 ? &Ae[(size_t)r * nd + d0] : &Ae[(size_t)d0 * nd + r];
                  C *b = axis == 0 ? &Ae[(size_t)r * nd + d0 + 1]
                                   : &Ae[(size_t)(d0 + 1) * nd + r];
                  const C na = Ti[0] * *a + Ti[1] * *b;
                  const C nb = Ti[2] * *a + Ti[3] * *b;
                  *a = na;
                  *b = nb;
               }
            }
         }
      }
   }
}

void orc_reassemble_element(const orc_asm_data *ad, const orc_problem *pb,
                            int e, int mask, int apply_te, double *Ae_out)
{
   const int nd = ad->nd;
   double *pc = (double *)malloc(sizeof(double) * 3 * nd);
   double *pv = (double *)malloc(sizeof(double) * 3 * nd);
   reassemble_element_eps(ad, pb, e, mask, apply_te, NULL, (C *)Ae_out, pc, pv);
   free(pc);
   free(pv);
}

void orc_reassemble_range(const orc_asm_data *ad, const orc_problem *pb,
                          int e0, int e1, int mask, int apply_te, double *Ae_out)
{
   const int nd = ad->nd;
   double *pc = (double *)malloc(sizeof(double) * 3 * nd);
   double *pv = (double *)malloc(sizeof(double) * 3 * nd);
   for (int e = e0; e < e1; ++e)
   {
      reassemble_element_eps(ad, pb, e, mask, apply_te, NULL,
                             (C *)(Ae_out + (size_t)(e - e0) * nd * nd * 2), pc, pv);
   }
   free(pc);
   free(pv);
}

static int lower_bound_pos(const int *jb, const int *je, int c)
{
   const int *lo = jb;
   long n = je - jb;
   while (n > 0)
   {
      long h = n / 2;
      if (lo[h] < c) { lo += h + 1; n -= h + 1; }
      else { n = h; }
   }
   return (int)(lo - jb);
}

/* scatter of one element (solver/asm_data.inc:384-400), rows in [r0, r1) */
static int scatter_one(const orc_asm_data *ad, int e, const C *Ae, int r0, int r1, double *vre,
                       double *vim)
{
   const int nd = ad->nd;
   const int *ed = ad->dof + (size_t)e * nd;
   const double *es = ad->sign + (size_t)e * nd;
   for (int a = 0; a < nd; ++a)
   {
      const int row = ed[a];
      if (row < r0 || row >= r1) { continue; }
      const int *jb = ad->Jcol + ad->I[row];
      const int *je = ad->Jcol + ad->I[row + 1];
      for (int b = 0; b < nd; ++b)
      {
         const int off = lower_bound_pos(jb, je, ed[b]);
         if (jb + off == je || jb[off] != ed[b]) { return -1; }
         const size_t k = (size_t)ad->I[row] + off;
         const C v = es[a] * es[b] * Ae[(size_t)a * nd + b];
         vre[k] += creal(v);
         vim[k] += cimag(v);
      }
   }
   return 0;
}

int orc_reassemble_all(const orc_asm_data *ad, const orc_problem *pb,
                       double *vre, double *vim)
{
   const int nd = ad->nd;
   const size_t nnz = (size_t)ad->I[ad->height];
   memset(vre, 0, nnz * sizeof(double));
   memset(vim, 0, nnz * sizeof(double));
   C *Ae = (C *)malloc(sizeof(C) * nd * nd);
   double *pc = (double *)malloc(sizeof(double) * 3 * nd);
   double *pv = (double *)malloc(sizeof(double) * 3 * nd);
   const int mask = ORC_CURLCURL | ORC_MASS |
                    (ad->has_kpar ? (ORC_CROSS0 | ORC_CROSS1 | ORC_KLK) : 0);
   int rc = 0;
   for (int e = 0; e < ad->ne && rc == 0; ++e)
   {
      reassemble_element_eps(ad, pb, e, mask, 1, NULL, Ae, pc, pv);
      rc = scatter_one(ad, e, Ae, 0, ad->height, vre, vim);
   }
   free(Ae);
   free(pc);
   free(pv);
   return rc;
}

int orc_scatter(const orc_asm_data *ad, const double *Ae_all, double *vre, double *vim)
{
   const size_t nnz = (size_t)ad->I[ad->height];
   const size_t n2 = (size_t)ad->nd * ad->nd;
   memset(vre, 0, nnz * sizeof(double));
   memset(vim, 0, nnz * sizeof(double));
   for (int e = 0; e < ad->ne; ++e)
   {
      if (scatter_one(ad, e, (const C *)(Ae_all + (size_t)e * n2 * 2), 0, ad->height, vre, vim))
      {
         return -1;
      }
   }
   return 0;
}

int orc_scatter_elements(const orc_asm_data *ad, const double *Ae_range, int e0, int e1,
                         double *vre, double *vim)
{
   const size_t n2 = (size_t)ad->nd * ad->nd;
   for (int e = e0; e < e1; ++e)
   {
      if (scatter_one(ad, e, (const C *)(Ae_range + (size_t)(e - e0) * n2 * 2), 0, ad->height,
                      vre, vim))
      {
         return -1;
      }
   }
   return 0;
}

int orc_scatter_rows(const orc_asm_data *ad, const double *Ae_all, int r0, int r1, double *vre,
                     double *vim)
{
   const size_t k0 = (size_t)ad->I[r0], k1 = (size_t)ad->I[r1];
   const size_t n2 = (size_t)ad->nd * ad->nd;
   memset(vre + k0, 0, (k1 - k0) * sizeof(double));
   memset(vim + k0, 0, (k1 - k0) * sizeof(double));
   for (int e = 0; e < ad->ne; ++e)
   {
      if (scatter_one(ad, e, (const C *)(Ae_all + (size_t)e * n2 * 2), r0, r1, vre, vim))
      {
         return -1;
      }
   }
   return 0;
}

static int cmp_int(const void *a, const void *b)
{
   const int x = *(const int *)a, y = *(const int *)b;
   return (x > y) - (x < y);
}

/* solver/asm_data.inc:178-210: per-row column lists from every element's
 * dof x dof block, then sort + unique per row. */
int64_t orc_build_csr_symbolic(int ne, int nd, int height, const int *dof,
                               int *I_out, int **Jcol_out)
{
   int64_t *cnt = (int64_t *)calloc((size_t)height + 1, sizeof(int64_t));
   if (!cnt) { return -1; }
   for (size_t i = 0; i < (size_t)ne * nd; ++i) { cnt[dof[i] + 1] += nd; }
   for (int r = 0; r < height; ++r) { cnt[r + 1] += cnt[r]; }
   int *cols = (int *)malloc(sizeof(int) * (size_t)cnt[height]);
   int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * ((size_t)height + 1));
   if (!cols || !fill) { free(cnt); free(cols); free(fill); return -1; }
   memcpy(fill, cnt, sizeof(int64_t) * ((size_t)height + 1));
   for (int e = 0; e < ne; ++e)
   {
      const int *ed = dof + (size_t)e * nd;
      for (int a = 0; a < nd; ++a)
      {
         int *row = cols + fill[ed[a]];
         for (int b = 0; b < nd; ++b) { row[b] = ed[b]; }
         fill[ed[a]] += nd;
      }
   }
   int64_t nnz = 0;
   I_out[0] = 0;
   for (int r = 0; r < height; ++r)
   {
      int *row = cols + cnt[r];
      const int64_t len = cnt[r + 1] - cnt[r];
      qsort(row, (size_t)len, sizeof(int), cmp_int);
      int64_t u = 0;
      for (int64_t i = 0; i < len; ++i)
      {
         if (u == 0 || row[u - 1] != row[i]) { row[u++] = row[i]; }
      }
      fill[r] = u;
      nnz += u;
      I_out[r + 1] = (int)nnz;
   }
   int *Jc = (int *)malloc(sizeof(int) * (size_t)(nnz > 0 ? nnz : 1));
   for (int r = 0; r < height; ++r)
   {
      memcpy(Jc + I_out[r], cols + cnt[r], sizeof(int) * (size_t)fill[r]);
   }
   free(cols);
   free(cnt);
   free(fill);
   *Jcol_out = Jc;
   return nnz;
}

void orc_free(void *p) { free(p); }

/* solver/asm_gpu.inc:463-478 */
void orc_eliminate_rows(int r0, int r1, const int *Ip, const int *Jcol,
                        const unsigned char *is_ess, double *vre, double *vim)
{
   for (int r = r0; r < r1; ++r)
   {
      const int row_ess = is_ess[r];
      for (int k = Ip[r]; k < Ip[r + 1]; ++k)
      {
         const int c = Jcol[k];
         if (row_ess || is_ess[c])
         {
            vre[k] = (row_ess && c == r) ? 1.0 : 0.0;
            vim[k] = 0.0;
         }
      }
   }
}

void orc_eliminate(int n, const int *Ip, const int *Jcol,
                   const unsigned char *is_ess, double *vre, double *vim)
{
   orc_eliminate_rows(0, n, Ip, Jcol, is_ess, vre, vim);
}

/* solver/asm_gpu.inc:481-493 (row-wise complex SpMV) and :556-564 (norms) */
double orc_residual(int n, const int *Ip, const int *Jcol, const double *vre,
                    const double *vim, const double *X2n, const double *B2n)
{
   double rn = 0.0, bn = 0.0;
   for (int r = 0; r < n; ++r)
   {
      double yr = 0.0, yi = 0.0;
      for (int k = Ip[r]; k < Ip[r + 1]; ++k)
      {
         const int c = Jcol[k];
         const double xr = X2n[c], xi = X2n[n + c];
         yr = yr + (vre[k] * xr - vim[k] * xi);
         yi = yi + (vre[k] * xi + vim[k] * xr);
      }
      const double dr = yr - B2n[r], di = yi - B2n[n + r];
      rn += dr * dr + di * di;
      bn += B2n[r] * B2n[r] + B2n[n + r] * B2n[n + r];
   }
   return bn > 0.0 ? sqrt(rn / bn) : 0.0;
}

/* solver/problem_common.inc:219-236: incident envelope of one part */
static void incident_at(const orc_incident *inc, int rhs, int slab, double z, C e[3])
{
   const double *w = inc->waves + 18 * ((size_t)rhs * inc->nslab + slab);
   const C qA = ld(w + 2), qB = ld(w + 4);
   const C up = cexp(IU * qA * (z - w[0]));
   const C dn = cexp(IU * qB * (z - w[1]));
   for (int i = 0; i < 3; ++i) { e[i] = ld(w + 6 + 2 * i) * up + ld(w + 12 + 2 * i) * dn; }
}

static void face_dual_vector(const orc_asm_data *ad, int e, C *b)
{
   if (ad->p < 2) { return; }
   const int p = ad->p, foff = 6 * p, ntd = p * (p - 1);
   const int *fo = ad->fo + (size_t)e * 4;
   for (int f = 0; f < 4; ++f)
   {
      const double *Ti = orc_TInv + 4 * fo[f];
      for (int i = 0; i < ntd / 2; ++i)
      {
         const int d0 = foff + f * ntd + 2 * i;
         const C a = b[d0], c = b[d0 + 1];
         b[d0] = Ti[0] * a + Ti[1] * c;
         b[d0 + 1] = Ti[2] * a + Ti[3] * c;
      }
   }
}

/* ScatterSource (solver/problem_common.inc:238-266) integrated against the
 * physical shapes on the mass rule, then TransformDual of the element
 * vector. */
void orc_rhs_element(const orc_asm_data *ad, const orc_problem *pb, const orc_incident *inc,
                     int e, int rhs, int unit_deps, double *b_out)
{
   const int nd = ad->nd;
   C *b = (C *)b_out;
   for (int k = 0; k < nd; ++k) { b[k] = 0.0; }
   const int reg = ad->attr[e] - 1;
   /* source strength eps - eps_bg: the region's own (frustum regions,
    * solver/problem_common.inc:313-317) or, with per-element eps, that of
    * the element against its slab background eps_bg = eps_reg - deps_reg
    * (deps_reg = 0 outside source regions) -- so an element of a background
    * region whose eps differs from the slab's becomes a source as well */
   C de = inc->region_src[reg] ? ld(inc->region_deps + 2 * reg) : 0.0;
   if (ad->eps_elem) { de = ld(ad->eps_elem + 2 * (size_t)e) - (ld(pb->eps + 2 * reg) - de); }
   if (unit_deps) { de = 1.0; }
   if (de == 0.0) { return; }
   const double *J = ad->J + (size_t)e * 9;
   const double *Ji = ad->Jinv + (size_t)e * 9;
   const double detJ = ad->detJ[e];
   const double *v0 = ad->v0 + (size_t)e * 3;
   const C fac = pb->k0 * pb->k0 * de;
   const orc_rule *rm = &ad->rule_mass;
   for (int q = 0; q < rm->nq; ++q)
   {
      const double *xi = rm->ip + (size_t)q * 3;
      const double z = v0[2] + J[6] * xi[0] + J[7] * xi[1] + J[8] * xi[2];
      C ei[3];
      incident_at(inc, rhs, inc->region_slab[reg], z, ei);
      const double w = rm->w[q] * detJ;
      const double *vs = rm->vshape + (size_t)q * nd * 3;
      for (int k = 0; k < nd; ++k)
      {
         C s = 0.0;
         for (int i = 0; i < 3; ++i)
         {
            const double v = vs[k * 3 + 0] * Ji[0 * 3 + i] + vs[k * 3 + 1] * Ji[1 * 3 + i] +
                             vs[k * 3 + 2] * Ji[2 * 3 + i];
            s += v * (fac * ei[i]);
         }
         b[k] += w * s;
      }
   }
   face_dual_vector(ad, e, b);
}

void orc_rhs_all(const orc_asm_data *ad, const orc_problem *pb, const orc_incident *inc,
                 int rhs, double *b_out)
{
   const int nd = ad->nd;
   C *b = (C *)b_out;
   for (int i = 0; i < ad->height; ++i) { b[i] = 0.0; }
   C *be = (C *)malloc(sizeof(C) * nd);
   for (int e = 0; e < ad->ne; ++e)
   {
      orc_rhs_element(ad, pb, inc, e, rhs, 0, (double *)be);
      const int *ed = ad->dof + (size_t)e * nd;
      const double *es = ad->sign + (size_t)e * nd;
      for (int a = 0; a < nd; ++a) { b[ed[a]] += es[a] * be[a]; }
   }
   free(be);
}

/* Per-element sensitivity lam^T (dA/d eps_e u - db/d eps_e): the mass term
 * is linear in eps, so dA_e/d eps_e is the TE-transformed mass element
 * matrix with eps = 1; db/d eps_e is the element RHS at unit eps - eps_bg;
 * local vectors are sign * gathered global entries. */
void orc_eps_sensitivity_list(const orc_asm_data *ad, const orc_problem *pb,
                              const orc_incident *inc, int rhs_col, const double *lam,
                              const double *u, int nel, const int *elems, double *g)
{
   const int nd = ad->nd;
   C *Ae = (C *)malloc(sizeof(C) * nd * nd);
   C *be = (C *)malloc(sizeof(C) * nd);
   double *pc = (double *)malloc(sizeof(double) * 3 * nd);
   double *pv = (double *)malloc(sizeof(double) * 3 * nd);
   const C one = 1.0;
   for (int i = 0; i < nel; ++i)
   {
      const int e = elems ? elems[i] : i;
      reassemble_element_eps(ad, pb, e, ORC_MASS, 1, &one, Ae, pc, pv);
      const int use_src = inc && rhs_col >= 0;
      if (use_src) { orc_rhs_element(ad, pb, inc, e, rhs_col, 1, (double *)be); }
      const int *ed = ad->dof + (size_t)e * nd;
      const double *es = ad->sign + (size_t)e * nd;
      C acc = 0.0;
      for (int a = 0; a < nd; ++a)
      {
         const C la = es[a] * ld(lam + 2 * (size_t)ed[a]);
         C row = 0.0;
         for (int b = 0; b < nd; ++b)
         {
            row += Ae[(size_t)a * nd + b] * (es[b] * ld(u + 2 * (size_t)ed[b]));
         }
         if (use_src) { row -= be[a]; }
         acc += la * row;
      }
      g[2 * (size_t)i] = creal(acc);
      g[2 * (size_t)i + 1] = cimag(acc);
   }
   free(Ae);
   free(be);
   free(pc);
   free(pv);
}

void orc_eps_sensitivity(const orc_asm_data *ad, const orc_problem *pb,
                         const orc_incident *inc, int rhs_col,
                         const double *lam, const double *u, double *g)
{
   orc_eps_sensitivity_list(ad, pb, inc, rhs_col, lam, u, ad->ne, NULL, g);
}

/* ---- componentwise error scale -------------------------------------------
 * The scale of an entry is the value of its own expression tree evaluated on
 * the moduli of every factor, with every sum of terms a sum of their moduli
 * (no cancellation): s >= sum |terms|.  Any evaluation order of the entry in
 * floating point differs from the exact value by at most about
 * (number of operations) x 2^-53 x s (the running error bound of the
 * summation), so |g - r| / s compares two implementations entry by entry
 * independently of how much the entry cancels. */
static void element_scale(const orc_asm_data *ad, const orc_problem *pb, int e, int mask,
                          int apply_te, double *S, double *pc, double *pv)
{
   const int nd = ad->nd;
   const double *J = ad->J + (size_t)e * 9;
   const double *Ji = ad->Jinv + (size_t)e * 9;
   const double adet = fabs(ad->detJ[e]);
   const double *v0 = ad->v0 + (size_t)e * 3;
   const double aeps = ad->eps_elem ? cabs(ld(ad->eps_elem + 2 * (size_t)e))
                                    : cabs(ld(pb->eps + 2 * (ad->attr[e] - 1)));
   const double k2 = pb->k0 * pb->k0;
   const double aK[9] = {0.0, 0.0, fabs(pb->ky), 0.0, 0.0, fabs(pb->kx),
                         fabs(pb->ky), fabs(pb->kx), 0.0};
   for (size_t i = 0; i < (size_t)nd * nd; ++i) { S[i] = 0.0; }
   if (mask & ORC_CURLCURL)
   {
      const orc_rule *rc = &ad->rule_cc;
      for (int q = 0; q < rc->nq; ++q)
      {
         const double *xi = rc->ip + (size_t)q * 3;
         const double z = v0[2] + J[6] * xi[0] + J[7] * xi[1] + J[8] * xi[2];
         const C s = s_z(pb, z);
         const double li[3] = {1.0 / cabs(s), 1.0 / cabs(s), cabs(s)};
         const double w = fabs(rc->w[q]) * adet;
         const double *cs = rc->curl + (size_t)q * nd * 3;
         for (int k = 0; k < nd; ++k)
            for (int i = 0; i < 3; ++i)
            {
               pc[(size_t)k * 3 + i] = (fabs(cs[k * 3 + 0] * J[i * 3 + 0]) +
                                        fabs(cs[k * 3 + 1] * J[i * 3 + 1]) +
                                        fabs(cs[k * 3 + 2] * J[i * 3 + 2])) / adet;
            }
         for (int k = 0; k < nd; ++k)
            for (int l = 0; l < nd; ++l)
            {
               double t = 0.0;
               for (int i = 0; i < 3; ++i) { t += pc[(size_t)k * 3 + i] * li[i] * pc[(size_t)l * 3 + i]; }
               S[(size_t)k * nd + l] += w * t;
            }
      }
   }
   if (mask & (ORC_MASS | ORC_CROSS0 | ORC_CROSS1 | ORC_KLK))
   {
      const orc_rule *rm = &ad->rule_mass;
      for (int q = 0; q < rm->nq; ++q)
      {
         const double *xi = rm->ip + (size_t)q * 3;
         const double z = v0[2] + J[6] * xi[0] + J[7] * xi[1] + J[8] * xi[2];
         const C s = s_z(pb, z);
         const double as = cabs(s);
         const double lam[3] = {as, as, 1.0 / as};
         const double li[3] = {1.0 / as, 1.0 / as, as};
         const double w = fabs(rm->w[q]) * adet;
         const double *vs = rm->vshape + (size_t)q * nd * 3;
         const double *cs = rm->curl + (size_t)q * nd * 3;
         for (int k = 0; k < nd; ++k)
            for (int i = 0; i < 3; ++i)
            {
               pv[(size_t)k * 3 + i] = fabs(vs[k * 3 + 0] * Ji[0 * 3 + i]) +
                                       fabs(vs[k * 3 + 1] * Ji[1 * 3 + i]) +
                                       fabs(vs[k * 3 + 2] * Ji[2 * 3 + i]);
               pc[(size_t)k * 3 + i] = (fabs(cs[k * 3 + 0] * J[i * 3 + 0]) +
                                        fabs(cs[k * 3 + 1] * J[i * 3 + 1]) +
                                        fabs(cs[k * 3 + 2] * J[i * 3 + 2])) / adet;
            }
         double klk[9];
         for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c)
            {
               klk[r * 3 + c] = 0.0;
               for (int t = 0; t < 3; ++t) { klk[r * 3 + c] += aK[r * 3 + t] * li[t] * aK[t * 3 + c]; }
            }
         for (int k = 0; k < nd; ++k)
            for (int l = 0; l < nd; ++l)
            {
               double t = 0.0;
               if (mask & ORC_MASS)
               {
                  for (int i = 0; i < 3; ++i)
                  {
                     t += pv[(size_t)k * 3 + i] * (k2 * aeps * lam[i]) * pv[(size_t)l * 3 + i];
                  }
               }
               if (ad->has_kpar)
               {
                  for (int r = 0; r < 3; ++r)
                     for (int c = 0; c < 3; ++c)
                     {
                        if (mask & ORC_CROSS0)
                        {
                           t += pv[(size_t)k * 3 + r] * aK[r * 3 + c] * li[c] * pc[(size_t)l * 3 + c];
                        }
                        if (mask & ORC_CROSS1)
                        {
                           t += pc[(size_t)k * 3 + r] * li[r] * aK[r * 3 + c] * pv[(size_t)l * 3 + c];
                        }
                        if (mask & ORC_KLK)
                        {
                           t += pv[(size_t)k * 3 + r] * klk[r * 3 + c] * pv[(size_t)l * 3 + c];
                        }
                     }
               }
               S[(size_t)k * nd + l] += w * t;
            }
      }
   }
   if (apply_te && ad->p >= 2)
   {
      const int p = ad->p, foff = 6 * p, ntd = p * (p - 1);
      const int *fo = ad->fo + (size_t)e * 4;
      for (int axis = 0; axis < 2; ++axis)
         for (int f = 0; f < 4; ++f)
         {
            const double *Ti = orc_TInv + 4 * fo[f];
            for (int i = 0; i < ntd / 2; ++i)
            {
               const int d0 = foff + f * ntd + 2 * i;
               for (int r = 0; r < nd; ++r)
               {
                  double *a = axis == 0 ? &S[(size_t)r * nd + d0] : &S[(size_t)d0 * nd + r];
                  double *b = axis == 0 ? &S[(size_t)r * nd + d0 + 1] : &S[(size_t)(d0 + 1) * nd + r];
                  const double na = fabs(Ti[0]) * *a + fabs(Ti[1]) * *b;
                  const double nb = fabs(Ti[2]) * *a + fabs(Ti[3]) * *b;
                  *a = na;
                  *b = nb;
               }
            }
         }
   }
}

void orc_scale_range(const orc_asm_data *ad, const orc_problem *pb, int e0, int e1, int mask,
                     int apply_te, double *S_out)
{
   const int nd = ad->nd;
   double *pc = (double *)malloc(sizeof(double) * 3 * nd);
   double *pv = (double *)malloc(sizeof(double) * 3 * nd);
   for (int e = e0; e < e1; ++e)
   {
      element_scale(ad, pb, e, mask, apply_te, S_out + (size_t)(e - e0) * nd * nd, pc, pv);
   }
   free(pc);
   free(pv);
}

int orc_scale_scatter_rows(const orc_asm_data *ad, const double *S_all, int r0, int r1,
                           double *s)
{
   const int nd = ad->nd;
   const size_t k0 = (size_t)ad->I[r0], k1 = (size_t)ad->I[r1];
   memset(s + k0, 0, (k1 - k0) * sizeof(double));
   for (int e = 0; e < ad->ne; ++e)
   {
      const int *ed = ad->dof + (size_t)e * nd;
      const double *Se = S_all + (size_t)e * nd * nd;
      for (int a = 0; a < nd; ++a)
      {
         const int row = ed[a];
         if (row < r0 || row >= r1) { continue; }
         const int *jb = ad->Jcol + ad->I[row];
         const int *je = ad->Jcol + ad->I[row + 1];
         for (int b = 0; b < nd; ++b)
         {
            const int off = lower_bound_pos(jb, je, ed[b]);
            if (jb + off == je || jb[off] != ed[b]) { return -1; }
            s[(size_t)ad->I[row] + off] += Se[(size_t)a * nd + b];
         }
      }
   }
   return 0;
}

void orc_rhs_scale_all(const orc_asm_data *ad, const orc_problem *pb, const orc_incident *inc,
                       int rhs, double *s)
{
   const int nd = ad->nd;
   for (int i = 0; i < ad->height; ++i) { s[i] = 0.0; }
   double *be = (double *)malloc(sizeof(double) * nd);
   for (int e = 0; e < ad->ne; ++e)
   {
      const int reg = ad->attr[e] - 1;
      C de = inc->region_src[reg] ? ld(inc->region_deps + 2 * reg) : 0.0;
      if (ad->eps_elem) { de = ld(ad->eps_elem + 2 * (size_t)e) - (ld(pb->eps + 2 * reg) - de); }
      if (de == 0.0) { continue; }
      const double *J = ad->J + (size_t)e * 9;
      const double *Ji = ad->Jinv + (size_t)e * 9;
      const double adet = fabs(ad->detJ[e]);
      const double *v0 = ad->v0 + (size_t)e * 3;
      const double fac = pb->k0 * pb->k0 * cabs(de);
      const orc_rule *rm = &ad->rule_mass;
      for (int k = 0; k < nd; ++k) { be[k] = 0.0; }
      for (int q = 0; q < rm->nq; ++q)
      {
         const double *xi = rm->ip + (size_t)q * 3;
         const double z = v0[2] + J[6] * xi[0] + J[7] * xi[1] + J[8] * xi[2];
         const double *w = inc->waves + 18 * ((size_t)rhs * inc->nslab + inc->region_slab[reg]);
         const double up = cabs(cexp(IU * ld(w + 2) * (z - w[0])));
         const double dn = cabs(cexp(IU * ld(w + 4) * (z - w[1])));
         double ae[3];
         for (int i = 0; i < 3; ++i) { ae[i] = cabs(ld(w + 6 + 2 * i)) * up + cabs(ld(w + 12 + 2 * i)) * dn; }
         const double wq = fabs(rm->w[q]) * adet;
         const double *vs = rm->vshape + (size_t)q * nd * 3;
         for (int k = 0; k < nd; ++k)
         {
            double t = 0.0;
            for (int i = 0; i < 3; ++i)
            {
               const double v = fabs(vs[k * 3 + 0] * Ji[0 * 3 + i]) + fabs(vs[k * 3 + 1] * Ji[1 * 3 + i]) +
                                fabs(vs[k * 3 + 2] * Ji[2 * 3 + i]);
               t += v * fac * ae[i];
            }
            be[k] += wq * t;
         }
      }
      if (ad->p >= 2)
      {
         const int p = ad->p, foff = 6 * p, ntd = p * (p - 1);
         const int *fo = ad->fo + (size_t)e * 4;
         for (int f = 0; f < 4; ++f)
         {
            const double *Ti = orc_TInv + 4 * fo[f];
            for (int i = 0; i < ntd / 2; ++i)
            {
               const int d0 = foff + f * ntd + 2 * i;
               const double a = be[d0], c = be[d0 + 1];
               be[d0] = fabs(Ti[0]) * a + fabs(Ti[1]) * c;
               be[d0 + 1] = fabs(Ti[2]) * a + fabs(Ti[3]) * c;
            }
         }
      }
      const int *ed = ad->dof + (size_t)e * nd;
      for (int a = 0; a < nd; ++a) { s[ed[a]] += be[a]; }
   }
   free(be);
}

/* Affine geometry of ne elements from their vertex coordinates (ne x 4 x 3),
 * in the operation order of the library's device extraction: J columns
 * x_{j+1} - x_0, detJ by cofactors of the first row, J^-1 = adj(J) / detJ,
 * v0 = x_0 (the quantities solver/asm_data.inc:152-171 reads from the
 * element transformation). */
void orc_geometry(int ne, const double *xyz, double *J, double *Jinv, double *detJ, double *v0)
{
   for (int e = 0; e < ne; ++e)
   {
      const double *x = xyz + (size_t)e * 12;
      double a[9];
      for (int i = 0; i < 3; ++i)
         for (int j = 0; j < 3; ++j) { a[i * 3 + j] = x[(j + 1) * 3 + i] - x[i]; }
      const double c00 = a[4] * a[8] - a[5] * a[7];
      const double c01 = a[5] * a[6] - a[3] * a[8];
      const double c02 = a[3] * a[7] - a[4] * a[6];
      const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
      double *iv = Jinv + (size_t)e * 9;
      iv[0] = c00 / det;
      iv[3] = c01 / det;
      iv[6] = c02 / det;
      iv[1] = (a[2] * a[7] - a[1] * a[8]) / det;
      iv[4] = (a[0] * a[8] - a[2] * a[6]) / det;
      iv[7] = (a[1] * a[6] - a[0] * a[7]) / det;
      iv[2] = (a[1] * a[5] - a[2] * a[4]) / det;
      iv[5] = (a[2] * a[3] - a[0] * a[5]) / det;
      iv[8] = (a[0] * a[4] - a[1] * a[3]) / det;
      for (int i = 0; i < 9; ++i) { J[(size_t)e * 9 + i] = a[i]; }
      detJ[e] = det;
      for (int d = 0; d < 3; ++d) { v0[(size_t)e * 3 + d] = x[d]; }
   }
}
