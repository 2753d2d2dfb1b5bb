/* Sanitizer driver of the CPU oracle (built with -fsanitize=address,undefined):
 * reads a small case in the layout tests/test_shim.py writes (meta.txt + raw
 * arrays) and runs every oracle entry point over it.  Exit status 0 when
 * every call completes; the sanitizers abort on the first invalid access or
 * undefined operation. */
#include "../../oracle/asm_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static const char *g_dir;

static void *read_file(const char *name, size_t *bytes)
{
   char path[4096];
   snprintf(path, sizeof(path), "%s/%s", g_dir, name);
   FILE *f = fopen(path, "rb");
   if (!f) { fprintf(stderr, "missing %s\n", path); exit(2); }
   fseek(f, 0, SEEK_END);
   const long n = ftell(f);
   fseek(f, 0, SEEK_SET);
   void *p = malloc(n > 0 ? (size_t)n : 1);
   if (n > 0 && fread(p, 1, (size_t)n, f) != (size_t)n) { exit(2); }
   fclose(f);
   if (bytes) { *bytes = (size_t)n; }
   return p;
}

static double meta(const char *key)
{
   char path[4096], k[128];
   double v;
   snprintf(path, sizeof(path), "%s/meta.txt", g_dir);
   FILE *f = fopen(path, "r");
   if (!f) { exit(2); }
   while (fscanf(f, "%127s %lf", k, &v) == 2)
   {
      if (strcmp(k, key) == 0) { fclose(f); return v; }
   }
   fclose(f);
   fprintf(stderr, "meta key %s missing\n", key);
   exit(2);
}

int main(int argc, char **argv)
{
   if (argc != 2) { return 2; }
   g_dir = argv[1];
   orc_asm_data ad;
   memset(&ad, 0, sizeof(ad));
   ad.ne = (int)meta("ne");
   ad.nd = (int)meta("nd");
   ad.p = (int)meta("p");
   ad.height = (int)meta("height");
   ad.has_kpar = meta("kx") != 0.0 || meta("ky") != 0.0;
   ad.dof = read_file("dof.i32", NULL);
   ad.sign = read_file("sign.f64", NULL);
   ad.fo = ad.p >= 2 ? read_file("fo.i32", NULL) : calloc((size_t)ad.ne * 4, sizeof(int));
   ad.attr = read_file("attr.i32", NULL);
   ad.J = read_file("J.f64", NULL);
   ad.Jinv = read_file("Jinv.f64", NULL);
   ad.detJ = read_file("detJ.f64", NULL);
   ad.v0 = read_file("v0.f64", NULL);
   const char *pre[2] = {"cc", "mass"};
   orc_rule *rules[2] = {&ad.rule_cc, &ad.rule_mass};
   for (int r = 0; r < 2; ++r)
   {
      char nm[64];
      size_t b;
      snprintf(nm, sizeof(nm), "%s_w.f64", pre[r]);
      rules[r]->w = read_file(nm, &b);
      rules[r]->nq = (int)(b / sizeof(double));
      snprintf(nm, sizeof(nm), "%s_ip.f64", pre[r]);
      rules[r]->ip = read_file(nm, NULL);
      snprintf(nm, sizeof(nm), "%s_vs.f64", pre[r]);
      rules[r]->vshape = read_file(nm, NULL);
      snprintf(nm, sizeof(nm), "%s_curl.f64", pre[r]);
      rules[r]->curl = read_file(nm, NULL);
   }
   ad.I = read_file("I.i32", NULL);
   ad.Jcol = read_file("Jcol.i32", NULL);
   size_t b;
   orc_problem pb;
   memset(&pb, 0, sizeof(pb));
   pb.k0 = meta("k0");
   pb.kx = meta("kx");
   pb.ky = meta("ky");
   pb.z_min = meta("z_min");
   pb.z_max = meta("z_max");
   pb.pml_thick = meta("pml_thick");
   pb.pml_order = meta("pml_order");
   pb.pml_target = meta("pml_target");
   pb.eps = read_file("eps.f64", &b);
   pb.nreg = (int)(b / (2 * sizeof(double)));
   pb.is_pml = read_file("is_pml.i32", NULL);
   pb.pml_top = read_file("pml_top.i32", NULL);

   const int ne = ad.ne, nd = ad.nd, n = ad.height;
   const size_t nnz = (size_t)ad.I[n];
   double *Ae = malloc(sizeof(double) * 2 * (size_t)ne * nd * nd);
   double *S = malloc(sizeof(double) * (size_t)ne * nd * nd);
   double *vre = malloc(sizeof(double) * nnz), *vim = malloc(sizeof(double) * nnz);
   double *s = malloc(sizeof(double) * nnz);
   orc_reassemble_range(&ad, &pb, 0, ne, ORC_CURLCURL | ORC_MASS | ORC_CROSS0 | ORC_CROSS1 | ORC_KLK,
                        1, Ae);
   orc_reassemble_element(&ad, &pb, ne - 1, ORC_KLK, 0, Ae);
   orc_reassemble_range(&ad, &pb, 0, ne, ORC_CURLCURL | ORC_MASS | ORC_CROSS0 | ORC_CROSS1 | ORC_KLK,
                        1, Ae);
   if (orc_reassemble_all(&ad, &pb, vre, vim) != 0) { return 3; }
   if (orc_scatter(&ad, Ae, vre, vim) != 0) { return 3; }
   if (orc_scatter_rows(&ad, Ae, n / 3, n, vre, vim) != 0) { return 3; }
   if (orc_scatter_elements(&ad, Ae, 0, ne, vre, vim) != 0) { return 3; }
   orc_scale_range(&ad, &pb, 0, ne, ORC_CURLCURL | ORC_MASS | ORC_CROSS0 | ORC_CROSS1 | ORC_KLK, 1, S);
   if (orc_scale_scatter_rows(&ad, S, 0, n, s) != 0) { return 3; }
   int *I2 = malloc(sizeof(int) * ((size_t)n + 1));
   int *J2 = NULL;
   const int64_t nnz2 = orc_build_csr_symbolic(ne, nd, n, ad.dof, I2, &J2);
   if (nnz2 != (int64_t)nnz) { return 4; }
   orc_free(J2);
   unsigned char *ess = calloc((size_t)n, 1);
   for (int i = 0; i < n; i += 7) { ess[i] = 1; }
   orc_eliminate(n, ad.I, ad.Jcol, ess, vre, vim);
   double *X = malloc(sizeof(double) * 2 * (size_t)n), *B = malloc(sizeof(double) * 2 * (size_t)n);
   for (int i = 0; i < 2 * n; ++i) { X[i] = 1.0 + 0.01 * (i % 13); B[i] = 0.5 - 0.02 * (i % 11); }
   (void)orc_residual(n, ad.I, ad.Jcol, vre, vim, X, B);
   double sz[2];
   orc_s_z(&pb, pb.z_max + 0.5 * pb.pml_thick, sz);
   /* a plane-wave incident field in every slab region, sources everywhere */
   const int nslab = 2;
   double *waves = calloc(18 * (size_t)nslab, sizeof(double));
   for (int sl = 0; sl < nslab; ++sl)
   {
      waves[18 * sl + 2] = 0.3;
      waves[18 * sl + 4] = -0.3;
      waves[18 * sl + 8] = 1.0;
   }
   int *rslab = calloc((size_t)pb.nreg, sizeof(int)), *rsrc = calloc((size_t)pb.nreg, sizeof(int));
   double *rdeps = calloc(2 * (size_t)pb.nreg, sizeof(double));
   for (int r = 0; r < pb.nreg; ++r) { rslab[r] = r % nslab; rsrc[r] = 1; rdeps[2 * r] = 0.1; }
   orc_incident inc = {1, nslab, waves, rslab, rdeps, rsrc};
   double *bv = malloc(sizeof(double) * 2 * (size_t)n), *bs = malloc(sizeof(double) * (size_t)n);
   orc_rhs_all(&ad, &pb, &inc, 0, bv);
   orc_rhs_scale_all(&ad, &pb, &inc, 0, bs);
   double *g = malloc(sizeof(double) * 2 * (size_t)ne);
   orc_eps_sensitivity(&ad, &pb, &inc, 0, X, B, g);
   orc_eps_sensitivity(&ad, &pb, NULL, -1, X, B, g);
   /* per-element eps on the same data */
   double *ee = malloc(sizeof(double) * 2 * (size_t)ne);
   for (int e = 0; e < ne; ++e) { ee[2 * e] = 1.0 + 0.001 * e; ee[2 * e + 1] = 0.01; }
   ad.eps_elem = ee;
   if (orc_reassemble_all(&ad, &pb, vre, vim) != 0) { return 3; }
   orc_rhs_all(&ad, &pb, &inc, 0, bv);
   /* geometry from vertex coordinates (a unit tet per element) */
   double *xyz = calloc(12 * (size_t)ne, sizeof(double));
   for (int e = 0; e < ne; ++e)
   {
      double *x = xyz + 12 * (size_t)e;
      x[0] = e;
      x[3] = e + 1.0;
      x[7] = 1.0;
      x[11] = 1.0;
   }
   double *Jg = malloc(sizeof(double) * 9 * (size_t)ne), *Jig = malloc(sizeof(double) * 9 * (size_t)ne);
   double *dg = malloc(sizeof(double) * (size_t)ne), *vg = malloc(sizeof(double) * 3 * (size_t)ne);
   orc_geometry(ne, xyz, Jg, Jig, dg, vg);
   printf("san_oracle ok ne=%d nd=%d n=%d nnz=%zu\n", ne, nd, n, nnz);
   return 0;
}
