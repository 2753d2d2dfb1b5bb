// Test driver of include/asm_hip.inc: the reference solve driver's assembly
// sequence (solver/lithofem_solve.cpp:783-792) through the drop-in header,
// with fixture types that carry the member names the reference's AsmData,
// AsmRefRule, Region, Problem, mfem::Array<int>, mfem::Vector,
// SparseMatrix / ComplexSparseMatrix and the UMFPACK solver have.
//
//   shim_driver <dir>
//
// reads the mesh arrays a test wrote into <dir> (meta.txt + raw arrays),
// runs Init -> BuildGlobal -> ComputeLocal -> ScatterEliminate, hands the
// system to the host (AsmHostSystem, TimedDirectSolve with a fixture solver
// that applies the matrix), computes the device residual, and writes
// vals_re.f64 / vals_im.f64 / y.f64 into <dir>.  Exit status 0 on success,
// 3 when the assembly fails (the driver would fall back to MFEM assembly).
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <string>
#include <vector>

struct TestRule   // AsmRefRule (solver/asm_data.inc:38-45)
{
   int nq = 0;
   std::vector<double> ip, w, vshape, curl;
};

struct TestAsmData   // AsmData (solver/asm_data.inc:77-97)
{
   int ne = 0, nd = 0, p = 0, height = 0;
   bool has_kpar = false;
   std::vector<int> dof;
   std::vector<double> sign;
   std::vector<int> fo, attr;
   std::vector<double> J, Jinv, detJ, v0;
   TestRule rule_cc, rule_mass;
   std::vector<int> I, Jcol;
};

struct TestRegion   // Region (solver/problem_common.inc:18-28)
{
   std::complex<double> eps;
   bool is_pml = false, pml_top = false;
};

struct TestProblem   // Problem (solver/problem_common.inc:40-65)
{
   double k0 = 0, kx = 0, ky = 0, z_min = 0, z_max = 0;
   double pml_order = 2, pml_target = 1e-8, pml_thick = 0;
   std::vector<TestRegion> regions;
};

struct TestIntArray   // mfem::Array<int>
{
   std::vector<int> v;
   int Size() const { return (int)v.size(); }
   int operator[](int i) const { return v[(size_t)i]; }
};

struct TestVector   // mfem::Vector
{
   std::vector<double> v;
   double *GetData() const { return const_cast<double *>(v.data()); }
};

struct TestSparse   // mfem::SparseMatrix(i, j, data, m, n, ownij, owna, issorted)
{
   const int *I, *J;
   const double *a;
   int m;
   TestSparse(int *i, int *j, double *d, int rows, int, bool, bool, bool) : I(i), J(j), a(d), m(rows) {}
};

struct TestComplexSparse   // mfem::ComplexSparseMatrix(re, im, own_re, own_im)
{
   const TestSparse *re, *im;
   TestComplexSparse(TestSparse *r, TestSparse *i, bool, bool) : re(r), im(i) {}
};

struct TestSolver   // the UMFPACK solver's SetOperator / Mult; Mult applies A
{
   const TestComplexSparse *A = nullptr;
   void SetOperator(const TestComplexSparse &op) { A = &op; }
   void Mult(const TestVector &x, TestVector &y) const
   {
      const int n = A->re->m;
      y.v.assign(2 * (size_t)n, 0.0);
      for (int r = 0; r < n; ++r)
         for (int k = A->re->I[r]; k < A->re->I[r + 1]; ++k)
         {
            const double ar = A->re->a[k], ai = A->im->a[k];
            const double xr = x.v[(size_t)A->re->J[k]], xi = x.v[(size_t)n + A->re->J[k]];
            y.v[(size_t)r] += ar * xr - ai * xi;
            y.v[(size_t)n + r] += ar * xi + ai * xr;
         }
   }
};

TestProblem prob;
double asm_nd_TInv[24];

#include "asm_hip.inc"

namespace {

std::string g_dir;

template <class T>
std::vector<T> read_array(const std::string &name)
{
   std::ifstream f(g_dir + "/" + name, std::ios::binary | std::ios::ate);
   if (!f) { std::fprintf(stderr, "missing %s\n", name.c_str()); std::exit(2); }
   const std::streamsize bytes = f.tellg();
   f.seekg(0);
   std::vector<T> v((size_t)bytes / sizeof(T));
   f.read(reinterpret_cast<char *>(v.data()), bytes);
   return v;
}

template <class T>
void write_array(const std::string &name, const std::vector<T> &v)
{
   std::ofstream f(g_dir + "/" + name, std::ios::binary);
   f.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
}

TestRule read_rule(const std::string &pre, int nq)
{
   TestRule r;
   r.nq = nq;
   r.ip = read_array<double>(pre + "_ip.f64");
   r.w = read_array<double>(pre + "_w.f64");
   r.vshape = read_array<double>(pre + "_vs.f64");
   r.curl = read_array<double>(pre + "_curl.f64");
   return r;
}

}  // namespace

int main(int argc, char **argv)
{
   if (argc != 2) { std::fprintf(stderr, "usage: shim_driver <dir>\n"); return 2; }
   g_dir = argv[1];
   std::map<std::string, double> meta;
   {
      std::ifstream f(g_dir + "/meta.txt");
      std::string k;
      double v;
      while (f >> k >> v) { meta[k] = v; }
   }
   TestAsmData ad;
   ad.ne = (int)meta["ne"];
   ad.nd = (int)meta["nd"];
   ad.p = (int)meta["p"];
   ad.height = (int)meta["height"];
   ad.dof = read_array<int>("dof.i32");
   ad.sign = read_array<double>("sign.f64");
   if (ad.p >= 2) { ad.fo = read_array<int>("fo.i32"); }
   ad.attr = read_array<int>("attr.i32");
   ad.J = read_array<double>("J.f64");
   ad.Jinv = read_array<double>("Jinv.f64");
   ad.detJ = read_array<double>("detJ.f64");
   ad.v0 = read_array<double>("v0.f64");
   ad.rule_cc = read_rule("cc", (int)meta["nq_cc"]);
   ad.rule_mass = read_rule("mass", (int)meta["nq_mass"]);
   ad.I = read_array<int>("I.i32");
   ad.Jcol = read_array<int>("Jcol.i32");
   prob.k0 = meta["k0"];
   prob.kx = meta["kx"];
   prob.ky = meta["ky"];
   prob.z_min = meta["z_min"];
   prob.z_max = meta["z_max"];
   prob.pml_order = meta["pml_order"];
   prob.pml_target = meta["pml_target"];
   prob.pml_thick = meta["pml_thick"];
   {
      const std::vector<double> eps = read_array<double>("eps.f64");
      const std::vector<int> is_pml = read_array<int>("is_pml.i32");
      const std::vector<int> pml_top = read_array<int>("pml_top.i32");
      for (size_t r = 0; r < is_pml.size(); ++r)
      {
         prob.regions.push_back({{eps[2 * r], eps[2 * r + 1]}, is_pml[r] != 0, pml_top[r] != 0});
      }
      const std::vector<double> t = read_array<double>("tinv.f64");
      for (int i = 0; i < 24; ++i) { asm_nd_TInv[i] = t[(size_t)i]; }
   }
   TestIntArray ess;
   ess.v = read_array<int>("ess.i32");

   // the driver's sequence (solver/lithofem_solve.cpp:783-792)
   AsmGpuGlobal asm_gpu;
   const bool kpar = prob.kx != 0.0 || prob.ky != 0.0;
   const int mm = ASM_MASS | (kpar ? ASM_CROSS0 | ASM_CROSS1 | ASM_KLK : 0);
   const bool ok = asm_gpu.Init(ad, 0) && asm_gpu.BuildGlobal(ad, ess) &&
                   asm_gpu.ComputeLocal(ASM_CURLCURL, mm, true) && asm_gpu.ScatterEliminate(true);
   if (!ok)
   {
      std::printf("asm_gpu: falling back to MFEM cpu assembly\n");
      return 3;
   }
   std::printf("asm_gpu: n %d nnz %d ne %d nd %d kernel_ms %.3f\n", asm_gpu.n(), asm_gpu.nnz(),
               asm_gpu.ne(), asm_gpu.nd(), asm_gpu.KernelMs());

   // host handoff (UMFPACK path) and a device residual of the same system
   AsmHostSystem<TestSparse, TestComplexSparse> host;
   if (!host.From(asm_gpu)) { return 3; }
   const int n = asm_gpu.n();
   TestVector X, Y;
   X.v.resize(2 * (size_t)n);
   for (int i = 0; i < 2 * n; ++i) { X.v[(size_t)i] = 1.0 + 0.001 * (i % 97); }
   TestSolver solver;
   DirectSolveTimes t;
   TimedDirectSolve(solver, t, *host.A, X, Y);
   double rel = -1.0;
   if (!asm_gpu.Residual(X, Y, rel)) { return 3; }
   std::printf("asm_gpu: residual of A X against the host product %.3e\n", rel);
   write_array("vals_re.f64", host.re);
   write_array("vals_im.f64", host.im);
   write_array("y.f64", Y.v);
   return 0;
}
