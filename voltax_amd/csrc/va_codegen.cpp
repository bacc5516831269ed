// Run-time code generation for large Verilog-A programs (see va_codegen.hpp).
#include "va_codegen.hpp"

#include <hip/hiprtc.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <sstream>
#include <vector>

#include "status.hpp"
#include "va_eval.hpp"

namespace vxm {

struct VAGenKernel {
  hipModule_t module = nullptr;
  hipFunction_t fn = nullptr;
};

bool va_codegen_for_all() {
  const char* v = std::getenv("VXM_VA_CODEGEN");
  return v && std::strcmp(v, "all") == 0;
}

namespace {

// Exact text of a double constant (hexadecimal floating literal).
std::string literal(double x) {
  if (std::isnan(x)) return "__builtin_nan(\"\")";
  if (std::isinf(x)) return x > 0 ? "__builtin_inf()" : "(-__builtin_inf())";
  char buf[64];
  std::snprintf(buf, sizeof buf, "%a", x);
  return std::string("(") + buf + ")";
}

// Dual arithmetic with the interpreter's rules (va_eval.hpp), one partial
// per unknown (or per seeded parameter).
const char* kMacros = R"(
#define VX __device__ __forceinline__
#define MAP(expr) _Pragma("unroll") for (int j = 0; j < M; ++j) r.d[j] = (expr)
)";

const char* kPreamble = R"(
struct D { double v; double d[M]; };
VX D cst(double x) { D r; r.v = x;
#pragma unroll
  for (int j = 0; j < M; ++j) r.d[j] = 0.0; return r; }
template <bool PS> VX D par(double x, int k, int pbase) {
  D r = cst(x); if (PS && k >= pbase && k - pbase < M) r.d[k - pbase] = 1.0; return r; }
template <bool PS> VX D unk(double x, int i) { D r = cst(x); if (!PS) r.d[i] = 1.0; return r; }
VX D add(const D& a, const D& b) { D r; r.v = a.v + b.v; MAP(a.d[j] + b.d[j]); return r; }
VX D sub(const D& a, const D& b) { D r; r.v = a.v - b.v; MAP(a.d[j] - b.d[j]); return r; }
VX D mul(const D& a, const D& b) { D r; r.v = a.v * b.v; MAP(a.d[j] * b.v + a.v * b.d[j]); return r; }
VX D dvd(const D& a, const D& b) { D r; r.v = a.v / b.v; MAP((a.d[j] - r.v * b.d[j]) / b.v); return r; }
VX D neg(const D& a) { D r; r.v = -a.v; MAP(-a.d[j]); return r; }
VX D exp_(const D& a) { D r; r.v = exp(a.v); MAP(r.v * a.d[j]); return r; }
VX D log_(const D& a) { D r; r.v = log(a.v); MAP(a.d[j] / a.v); return r; }
VX D sqrt_(const D& a) { D r; r.v = sqrt(a.v); MAP(a.d[j] / (2.0 * r.v)); return r; }
VX D abs_(const D& a) { D r; r.v = fabs(a.v); MAP(a.v < 0.0 ? -a.d[j] : a.d[j]); return r; }
VX D pow_(const D& a, const D& b) { D r; r.v = pow(a.v, b.v);
  const double da = b.v * pow(a.v, b.v - 1.0);
  MAP(da * a.d[j] + (b.d[j] != 0.0 ? r.v * log(a.v) * b.d[j] : 0.0)); return r; }
VX D min_(const D& a, const D& b) { return (a.v <= b.v) ? a : b; }
VX D max_(const D& a, const D& b) { return (a.v >= b.v) ? a : b; }
VX D limexp_(const D& a) { D r; const double e = exp(fmin(a.v, 80.0));
  r.v = a.v > 80.0 ? e * (1.0 + a.v - 80.0) : e; MAP(e * a.d[j]); return r; }
VX D tanh_(const D& a) { D r; r.v = tanh(a.v); MAP((1.0 - r.v * r.v) * a.d[j]); return r; }
VX D sin_(const D& a) { D r; r.v = sin(a.v); MAP(cos(a.v) * a.d[j]); return r; }
VX D cos_(const D& a) { D r; r.v = cos(a.v); MAP(-sin(a.v) * a.d[j]); return r; }
VX D flag(bool c) { return cst(c ? 1.0 : 0.0); }
VX D trunc_(const D& a) { return cst(trunc(a.v)); }
VX D round_(const D& a) { return cst(isfinite(a.v) ? copysign(floor(fabs(a.v) + 0.5), a.v) : a.v); }
VX D select_(const D& c, const D& a, const D& b) { return c.v != 0.0 ? a : b; }
VX D limit_(const D& c, const D& b, const D& g) { D r = b;
  const bool general = c.v != 0.0 && isfinite(g.v); if (general) { MAP(g.d[j]); } return r; }
)";

// One (instance, device) of the program: thread gid of its block range.
const char* kRun = R"(
VX void run(long gid, int nd, const int* __restrict__ devs, int B, const int* __restrict__ unk,
            int uw, const int* __restrict__ poff, const int* __restrict__ roff,
            const int* __restrict__ joff, int S, int P, int Cres, int Cjac,
            const double* __restrict__ params, const double* __restrict__ z,
            const double* __restrict__ lam, double cf, double cq,
            const double* __restrict__ scal, int mode, double scale,
            double* __restrict__ work, double* __restrict__ grad) {
  if (gid >= (long)B * nd) return;
  const int b = (int)(gid / nd), d = devs[gid % nd];
  if (scal) { cf = scal[1]; cq = scal[2]; }
  const double* zb = z + (long)b * S;
  double u[M], l[M];
#pragma unroll
  for (int i = 0; i < M; ++i) {
    const int idx = unk[(long)uw * d + i];
    u[i] = idx >= 0 ? zb[idx] : 0.0;
    l[i] = (lam && idx >= 0) ? lam[(long)b * S + idx] : 0.0;
  }
  const double* p = params + (long)b * P + poff[d];
  D F[M], Q[M];
  if (mode == 3) {
    double* gb = grad + (long)b * P + poff[d];
    for (int base = 0; base < NPAR; base += M) {
      body<true>(p, u, base, F, Q);
      for (int j = 0; j < M && base + j < NPAR; ++j) {
        double acc = 0.0;
        for (int i = 0; i < M; ++i) acc += l[i] * (cf * F[i].d[j] + cq * Q[i].d[j]);
        gb[base + j] += scale * acc;
      }
    }
    return;
  }
  body<false>(p, u, 0, F, Q);
  double* wb = work + (long)b * (Cres + Cjac);
  if (mode == 2) {
    for (int j = 0; j < M; ++j) {
      double acc = 0.0;
      for (int i = 0; i < M; ++i) acc += (cf * F[i].d[j] + cq * Q[i].d[j]) * l[i];
      wb[roff[d] + j] = acc;
    }
    return;
  }
  for (int i = 0; i < M; ++i) wb[roff[d] + i] = cf * F[i].v + cq * Q[i].v;
  if (mode == 1) {
    double* wj = wb + Cres + joff[d];
    for (int i = 0; i < M; ++i)
      for (int j = 0; j < M; ++j) wj[i * M + j] = cf * F[i].d[j] + cq * Q[i].d[j];
  }
}
)";

std::mutex cache_mu;
std::map<std::string, std::vector<char>> code_cache;  // source -> code object

int compile(const std::string& src, std::vector<char>& out) {
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "vxm_va_program.hip", 0, nullptr, nullptr) !=
      HIPRTC_SUCCESS)
    return VXM_ERR_HIP;
  hipDeviceProp_t prop;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
    (void)hiprtcDestroyProgram(&prog);
    return VXM_ERR_HIP;
  }
  std::string arch = std::string("--offload-arch=") + prop.gcnArchName;
  const char* opts[] = {arch.c_str(), "-O3", "-std=c++17", "-Xclang", "-target-feature",
                        "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult r = hiprtcCompileProgram(prog, 7, opts);
  if (r != HIPRTC_SUCCESS) {
    size_t n = 0;
    if (hiprtcGetProgramLogSize(prog, &n) == HIPRTC_SUCCESS && n > 1) {
      std::vector<char> log(n + 1, 0);
      if (hiprtcGetProgramLog(prog, log.data()) == HIPRTC_SUCCESS)
        std::fprintf(stderr, "voltax_mi355x: Verilog-A code generation failed:\n%s\n",
                     log.data());
    }
    (void)hiprtcDestroyProgram(&prog);
    return VXM_ERR_UNSUPPORTED;
  }
  size_t n = 0;
  if (hiprtcGetCodeSize(prog, &n) != HIPRTC_SUCCESS || n == 0) {
    (void)hiprtcDestroyProgram(&prog);
    return VXM_ERR_HIP;
  }
  out.assign(n, 0);
  const bool ok = hiprtcGetCode(prog, out.data()) == HIPRTC_SUCCESS;
  (void)hiprtcDestroyProgram(&prog);
  return ok ? VXM_OK : VXM_ERR_HIP;
}

}  // namespace

namespace {

// The code of program `prog` in namespace p<ns>: dual type, model body, run().
void program_source(std::ostringstream& o, const int* prog, int ns) {
  const int m = prog[1], nreg = prog[2], nins = prog[3], npar = prog[4];
  o << "// program " << ns << ": " << m << " unknowns, " << nins << " instructions\n";
  o << "namespace p" << ns << " {\n";
  o << "constexpr int M = " << m << ", NPAR = " << npar << ";\n" << kPreamble;
  o << "template <bool PS>\nVX void body(const double* __restrict__ p, const double* u, "
       "int pbase, D* F, D* Q) {\n";
  for (int r = 0; r < nreg; ++r) o << "  D r" << r << ";\n";
  const int* ins = prog + kVAHead;
  static const char* kBin[] = {nullptr, nullptr, nullptr, "add", "sub", "mul", "dvd"};
  for (int k = 0; k < nins; ++k, ins += kVAInsWords) {
    const int op = ins[0], dst = ins[1], a = ins[2], b = ins[3], c = ins[4];
    o << "  r" << dst << " = ";
    switch (op) {
      case VA_CONST: {  // the double's bits: low word a, high word b
        const uint64_t bits = ((uint64_t)(uint32_t)b << 32) | (uint32_t)a;
        double x;
        std::memcpy(&x, &bits, sizeof x);
        o << "cst(" << literal(x) << ")";
        break;
      }
      case VA_PARAM: o << "par<PS>(p[" << a << "], " << a << ", pbase)"; break;
      case VA_UNK: o << "unk<PS>(u[" << a << "], " << a << ")"; break;
      case VA_ADD: case VA_SUB: case VA_MUL: case VA_DIV:
        o << kBin[op] << "(r" << a << ", r" << b << ")";
        break;
      case VA_NEG: o << "neg(r" << a << ")"; break;
      case VA_EXP: o << "exp_(r" << a << ")"; break;
      case VA_LOG: o << "log_(r" << a << ")"; break;
      case VA_SQRT: o << "sqrt_(r" << a << ")"; break;
      case VA_ABS: o << "abs_(r" << a << ")"; break;
      case VA_POW: o << "pow_(r" << a << ", r" << b << ")"; break;
      case VA_MIN: o << "min_(r" << a << ", r" << b << ")"; break;
      case VA_MAX: o << "max_(r" << a << ", r" << b << ")"; break;
      case VA_LIMEXP: o << "limexp_(r" << a << ")"; break;
      case VA_TANH: o << "tanh_(r" << a << ")"; break;
      case VA_SIN: o << "sin_(r" << a << ")"; break;
      case VA_COS: o << "cos_(r" << a << ")"; break;
      case VA_LT: o << "flag(r" << a << ".v < r" << b << ".v)"; break;
      case VA_LE: o << "flag(r" << a << ".v <= r" << b << ".v)"; break;
      case VA_GT: o << "flag(r" << a << ".v > r" << b << ".v)"; break;
      case VA_GE: o << "flag(r" << a << ".v >= r" << b << ".v)"; break;
      case VA_EQ: o << "flag(r" << a << ".v == r" << b << ".v)"; break;
      case VA_NE: o << "flag(r" << a << ".v != r" << b << ".v)"; break;
      case VA_AND: o << "flag(r" << a << ".v != 0.0 && r" << b << ".v != 0.0)"; break;
      case VA_OR: o << "flag(r" << a << ".v != 0.0 || r" << b << ".v != 0.0)"; break;
      case VA_NOT: o << "flag(r" << a << ".v == 0.0)"; break;
      case VA_SELECT: o << "select_(r" << a << ", r" << b << ", r" << c << ")"; break;
      case VA_TRUNC: o << "trunc_(r" << a << ")"; break;
      case VA_ROUND: o << "round_(r" << a << ")"; break;
      case VA_LIMIT: o << "limit_(r" << a << ", r" << b << ", r" << c << ")"; break;
      default: o << "cst(0.0)"; break;
    }
    o << ";\n";
  }
  const int* out = prog + kVAHead + nins * kVAInsWords;
  for (int i = 0; i < m; ++i)
    o << "  F[" << i << "] = r" << out[i] << ";\n  Q[" << i << "] = r" << out[m + i] << ";\n";
  o << "}\n" << kRun << "}  // namespace p" << ns << "\n";
}

}  // namespace

std::string va_source(const int* const* progs, int n) {
  std::ostringstream o;
  o << "// generated from " << n << " voltax_amd Verilog-A program(s)\n" << kMacros;
  for (int g = 0; g < n; ++g) program_source(o, progs[g], g);
  // the dispatcher: block range g runs program g (VAParts mirrors VAGroupArgs)
  o << "struct VAPart { int nd, blk0; const int* devs; };\n"
       "struct VAParts { int n; VAPart g["
    << kVAGroup << "]; };\n"
       "extern \"C\" __global__ void __launch_bounds__(128)\n"
       "vxm_va_program(VAParts parts, int B, const int* __restrict__ unk, int uw,\n"
       "               const int* __restrict__ poff, const int* __restrict__ roff,\n"
       "               const int* __restrict__ joff, int S, int P, int Cres, int Cjac,\n"
       "               const double* __restrict__ params, const double* __restrict__ z,\n"
       "               const double* __restrict__ lam, double cf, double cq,\n"
       "               const double* __restrict__ scal, int mode, double scale,\n"
       "               double* __restrict__ work, double* __restrict__ grad) {\n"
       "  int g = 0;\n"
       "  for (int i = 1; i < parts.n; ++i)\n"
       "    if ((int)blockIdx.x >= parts.g[i].blk0) g = i;\n"
       "  const long gid = (long)((int)blockIdx.x - parts.g[g].blk0) * blockDim.x + threadIdx.x;\n"
       "  switch (g) {\n";
  for (int g = 0; g < n; ++g)
    o << "    case " << g << ": p" << g << "::run(gid, parts.g[" << g << "].nd, parts.g[" << g
      << "].devs, B, unk, uw, poff, roff, joff, S, P, Cres, Cjac, params, z, lam, cf, cq, "
         "scal, mode, scale, work, grad); break;\n";
  o << "  }\n}\n";
  return o.str();
}

int load_va_kernel(const int* const* progs, int n, VAGenKernel** out) {
  *out = nullptr;
  const std::string src = va_source(progs, n);
  std::vector<char> code;
  {
    std::lock_guard<std::mutex> lock(cache_mu);
    auto it = code_cache.find(src);
    if (it != code_cache.end()) {
      code = it->second;
    } else {
      if (int st = compile(src, code)) return st;
      code_cache.emplace(src, code);
    }
  }
  auto* k = new (std::nothrow) VAGenKernel();
  if (!k) return VXM_ERR_HOST_MEMORY;
  if (hipModuleLoadData(&k->module, code.data()) != hipSuccess) {
    delete k;
    return VXM_ERR_HIP;
  }
  if (hipModuleGetFunction(&k->fn, k->module, "vxm_va_program") != hipSuccess) {
    (void)hipModuleUnload(k->module);
    delete k;
    return VXM_ERR_HIP;
  }
  *out = k;
  return VXM_OK;
}

void release_va_kernels(Topology* t) {
  for (VAGenProgram& g : t->gen) {
    if (!g.kernel) continue;
    if (g.kernel->module) (void)hipModuleUnload(g.kernel->module);
    delete g.kernel;
    g.kernel = nullptr;
  }
}

int launch_va_generated(const TopoDev& t, int mode, int B, const double* params,
                        const double* z, const double* lam, double cf, double cq,
                        const double* scal, double scale, double* work, double* grad,
                        hipStream_t stream) {
  if (B <= 0) return VXM_OK;
  for (int g0 = 0; g0 < t.n_gen; ++g0) {
    const VAGenProgram& head = t.gen[g0];
    if (!head.kernel) continue;  // not the first program of a group
    VAGroupArgs parts{};
    parts.n = head.group_n;
    long blocks = 0;
    for (int i = 0; i < head.group_n; ++i) {
      const VAGenProgram& q = t.gen[g0 + i];
      parts.g[i].nd = q.n_devs;
      parts.g[i].blk0 = (int)blocks;
      parts.g[i].devs = t.va_devs + q.dev_off;
      blocks += ((long)B * q.n_devs + 127) / 128;
    }
    if (blocks == 0) continue;
    if (blocks > 0x7fffffffL) return VXM_ERR_INVALID;
    int uw = t.uw, S = t.S, P = t.P, Cres = t.Cres, Cjac = t.Cjac;
    void* args[] = {&parts,         &B,    (void*)&t.unk,  &uw,          (void*)&t.poff,
                    (void*)&t.roff, (void*)&t.joff, &S,    &P,           &Cres,
                    &Cjac,          (void*)&params, (void*)&z, (void*)&lam, &cf,
                    &cq,            (void*)&scal,   &mode, &scale,       &work,
                    &grad};
    if (hipModuleLaunchKernel(head.kernel->fn, (unsigned)blocks, 1, 1, 128, 1, 1, 0, stream,
                              args, nullptr) != hipSuccess)
      return VXM_ERR_HIP;
  }
  return VXM_OK;
}

}  // namespace vxm
