#include "topology.hpp"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <numeric>

#include "status.hpp"
#include "va_codegen.hpp"

namespace vxm {


Topology::~Topology() {
  release_va_kernels(this);
  if (dev_block) (void)hipFree(dev_block);
}

int build_topology(int S, int nv, int D, int P, int width, const int32_t* kind,
                   const int32_t* wave, const int32_t* unknowns, const int32_t* param_off,
                   Topology** out) {
  if (S <= 0 || nv < 0 || nv > S || D < 0 || P < 0 || !out) return VXM_ERR_INVALID;
  if (width < 4 || width > kMaxDeviceUnknowns) return VXM_ERR_INVALID;
  if (D > 0 && (!kind || !wave || !unknowns || !param_off)) return VXM_ERR_INVALID;
  auto* t = new Topology();
  t->S = S;
  t->nv = nv;
  t->D = D;
  t->P = P;
  t->kind.assign(kind, kind + D);
  t->wave.assign(wave, wave + D);
  t->uw = width;
  t->unk.assign(unknowns, unknowns + (size_t)width * D);
  t->poff.assign(param_off, param_off + D);
  t->nterm.resize(D);
  for (int d = 0; d < D; ++d) t->nterm[d] = kind_terminals(t->kind[d]);
  t->roff.resize(D);
  t->joff.resize(D);
  int cres = 0, cjac = 0;
  for (int d = 0; d < D; ++d) {
    int k = t->kind[d];
    if (k < KIND_R || k > KIND_VA16) {
      delete t;
      return VXM_ERR_INVALID;
    }
    int m = kind_m(k);
    if (m > width) {
      delete t;
      return VXM_ERR_INVALID;
    }
    for (int i = 0; i < width; ++i) {
      int u = t->unk[width * d + i];
      if (i >= m) {
        t->unk[width * d + i] = -1;
        continue;
      }
      if (u < -1 || u >= S) {
        delete t;
        return VXM_ERR_INVALID;
      }
    }
    if (t->poff[d] < 0 || t->poff[d] > P) {
      delete t;
      return VXM_ERR_INVALID;
    }
    t->roff[d] = cres;
    t->joff[d] = cjac;
    cres += m;
    cjac += m * m;
  }
  t->Cres = cres;
  t->Cjac = cjac;

  // --- CSR pattern: diagonal + every (row, col) pair of every device block
  std::vector<std::vector<int>> cols(S);
  for (int s = 0; s < S; ++s) cols[s].push_back(s);
  for (int d = 0; d < D; ++d) {
    int m = kind_m(t->kind[d]);
    for (int i = 0; i < m; ++i) {
      int r = t->unk[t->uw * d + i];
      if (r < 0) continue;
      for (int j = 0; j < m; ++j) {
        int c = t->unk[t->uw * d + j];
        if (c >= 0) cols[r].push_back(c);
      }
    }
  }
  t->indptr.assign(S + 1, 0);
  for (int s = 0; s < S; ++s) {
    auto& v = cols[s];
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    t->indptr[s + 1] = t->indptr[s] + (int)v.size();
  }
  t->nnz = t->indptr[S];
  t->indices.resize(t->nnz);
  for (int s = 0; s < S; ++s)
    std::copy(cols[s].begin(), cols[s].end(), t->indices.begin() + t->indptr[s]);
  auto slot = [&](int r, int c) -> int {
    auto b = t->indices.begin() + t->indptr[r], e = t->indices.begin() + t->indptr[r + 1];
    auto it = std::lower_bound(b, e, c);
    return (int)(it - t->indices.begin());
  };
  t->diag.resize(S);
  for (int s = 0; s < S; ++s) t->diag[s] = slot(s, s);

  // --- gather maps (device order, then local order)
  std::vector<int> rcount(S, 0), jcount(t->nnz, 0);
  for (int d = 0; d < D; ++d) {
    int m = kind_m(t->kind[d]);
    for (int i = 0; i < m; ++i) {
      int r = t->unk[t->uw * d + i];
      if (r < 0) continue;
      rcount[r]++;
      for (int j = 0; j < m; ++j) {
        int c = t->unk[t->uw * d + j];
        if (c >= 0) jcount[slot(r, c)]++;
      }
    }
  }
  t->rptr.assign(S + 1, 0);
  for (int s = 0; s < S; ++s) t->rptr[s + 1] = t->rptr[s] + rcount[s];
  t->jptr.assign(t->nnz + 1, 0);
  for (int p = 0; p < t->nnz; ++p) t->jptr[p + 1] = t->jptr[p] + jcount[p];
  t->ridx.resize(t->rptr[S]);
  t->jidx.resize(t->jptr[t->nnz]);
  std::vector<int> rfill(t->rptr.begin(), t->rptr.end() - 1);
  std::vector<int> jfill(t->jptr.begin(), t->jptr.end() - 1);
  for (int d = 0; d < D; ++d) {
    int m = kind_m(t->kind[d]);
    for (int i = 0; i < m; ++i) {
      int r = t->unk[t->uw * d + i];
      if (r < 0) continue;
      t->ridx[rfill[r]++] = t->roff[d] + i;
      for (int j = 0; j < m; ++j) {
        int c = t->unk[t->uw * d + j];
        if (c >= 0) t->jidx[jfill[slot(r, c)]++] = t->joff[d] + i * m + j;
      }
    }
  }

  t->slot_row.resize(t->nnz);
  for (int s = 0; s < S; ++s)
    for (int p = t->indptr[s]; p < t->indptr[s + 1]; ++p) t->slot_row[p] = s;
  for (int s = 0; s < S; ++s) {
    if (t->rptr[s + 1] - t->rptr[s] > kHeavyGather) t->heavy_r.push_back(s);
    if (t->indptr[s + 1] - t->indptr[s] > kHeavyGather) t->heavy_a.push_back(s);
  }
  for (int p = 0; p < t->nnz; ++p)
    if (t->jptr[p + 1] - t->jptr[p] > kHeavyGather) t->heavy_j.push_back(p);

  *out = t;
  return VXM_OK;
}

int upload_topology(Topology* t) {
  std::lock_guard<std::mutex> lock(t->mu);
  if (t->dev_block) return VXM_OK;
  // Verilog-A devices need their checked programs before any kernel runs
  for (int k : t->kind)
    if (kind_is_va(k) && !t->code_checked) return VXM_ERR_INVALID;
  // devices of generated programs are marked for the generated kernels only
  std::vector<int> kind_dev = t->kind;
  for (const VAGenProgram& g : t->gen)
    for (int i = 0; i < g.n_devs; ++i) kind_dev[t->va_devs[g.dev_off + i]] = KIND_VA_GEN;
  // --- upload all tables in one allocation
  std::vector<const std::vector<int>*> parts = {&kind_dev, &t->wave, &t->unk,  &t->poff,
                                                &t->roff, &t->joff, &t->indptr, &t->indices,
                                                &t->rptr, &t->ridx, &t->jptr,  &t->jidx,
                                                &t->diag, &t->code, &t->va_devs,
                                                &t->slot_row, &t->heavy_r, &t->heavy_j,
                                                &t->heavy_a};
  size_t total = 0;
  for (auto* p : parts) total += p->size() + 1;  // +1 keeps every pointer valid
  std::vector<int> host(total, 0);
  std::vector<size_t> offs;
  size_t o = 0;
  for (auto* p : parts) {
    offs.push_back(o);
    std::copy(p->begin(), p->end(), host.begin() + o);
    o += p->size() + 1;
  }
  void* block = nullptr;
  int ordinal = -1;
  if (hipGetDevice(&ordinal) != hipSuccess) return VXM_ERR_HIP;
  if (hipMalloc(&block, total * sizeof(int)) != hipSuccess) return VXM_ERR_HIP;
  if (hipMemcpy(block, host.data(), total * sizeof(int), hipMemcpyHostToDevice) !=
      hipSuccess) {
    (void)hipFree(block);
    return VXM_ERR_HIP;
  }
  t->dev_block = block;
  t->dev_ordinal = ordinal;
  const int* base = static_cast<const int*>(t->dev_block);
  TopoDev& v = t->dev;
  v.S = t->S;
  v.nv = t->nv;
  v.D = t->D;
  v.P = t->P;
  v.nnz = t->nnz;
  v.Cres = t->Cres;
  v.Cjac = t->Cjac;
  v.uw = t->uw;
  v.n_heavy_r = (int)t->heavy_r.size();
  v.n_heavy_j = (int)t->heavy_j.size();
  v.n_heavy_a = (int)t->heavy_a.size();
  const int** ptrs[] = {&v.kind, &v.wave, &v.unk,  &v.poff, &v.roff, &v.joff, &v.indptr,
                        &v.indices, &v.rptr, &v.ridx, &v.jptr, &v.jidx, &v.diag, &v.code,
                        &v.va_devs, &v.slot_row, &v.heavy_r, &v.heavy_j, &v.heavy_a};
  for (size_t i = 0; i < parts.size(); ++i) *ptrs[i] = base + offs[i];
  v.has_va = 0;
  for (int k : kind_dev) v.has_va |= kind_is_interpreted(k) ? 1 : 0;
  v.n_gen = (int)t->gen.size();
  v.gen = t->gen.data();
  // kernels of the generated programs, one per group of up to kVAGroup
  // (compiled once per source text per process)
  for (size_t g0 = 0; g0 < t->gen.size(); g0 += kVAGroup) {
    const int n = (int)std::min<size_t>(kVAGroup, t->gen.size() - g0);
    const int* progs[kVAGroup];
    for (int i = 0; i < n; ++i) progs[i] = t->code.data() + t->gen[g0 + i].code_off;
    if (int st = load_va_kernel(progs, n, &t->gen[g0].kernel)) {
      release_va_kernels(t);
      t->dev_block = nullptr;
      (void)hipFree(block);
      return st;
    }
    t->gen[g0].group_n = n;
  }
  return VXM_OK;
}

}  // namespace vxm

namespace vxm {

// Host check of the Verilog-A programs of every VA device (va_eval.hpp layout):
// every index the interpreter will use is in range, so the GPU never reads
// outside the program, the register file, the parameter row or u[].
int set_va_code(Topology* t, const int* words, int n_words) {
  std::lock_guard<std::mutex> lock(t->mu);
  if (t->dev_block) return VXM_ERR_INVALID;  // already uploaded
  if (n_words < 0 || (n_words > 0 && !words)) return VXM_ERR_INVALID;
  std::vector<int> code(words, words + n_words);
  constexpr int kMagic = 0x5642, kHead = 6, kIns = 5, kOps = 32;
  constexpr int kConst = 0, kParam = 1, kUnk = 2, kSelect = 28, kLimit = 31;
  const bool gen_all = va_codegen_for_all();
  std::vector<int> gen_of(t->D, -1);
  std::vector<VAGenProgram> gen;
  std::vector<int> offs_seen;
  for (int d = 0; d < t->D; ++d) {
    const int kind = t->kind[d];
    if (!kind_is_va(kind)) continue;
    const int64_t off = t->wave[d];
    if (off < 0 || off + kHead > n_words) return VXM_ERR_INVALID;
    const int* pr = code.data() + off;
    const int m = pr[1], nreg = pr[2], nins = pr[3], npar = pr[4], nterm = pr[5];
    // the interpreter holds 64 registers of 4 partials; larger programs run
    // as generated code (va_codegen.cpp), whose registers the compiler places
    const bool generated = gen_all || m > 4 || nreg > kInterpretedRegs;
    if (pr[0] != kMagic || m != kind_m(kind) || nreg < 1 || nreg > kGeneratedRegs ||
        nins < 0 || npar < 0 || t->poff[d] + npar > t->P || nterm < 1 || nterm > m)
      return VXM_ERR_INVALID;
    if (off + kHead + (int64_t)nins * kIns + 2 * m > n_words) return VXM_ERR_INVALID;
    for (int k = 0; k < nins; ++k) {
      const int* in = pr + kHead + k * kIns;
      const int op = in[0], dst = in[1];
      if (op < 0 || op >= kOps || dst < 0 || dst >= nreg) return VXM_ERR_INVALID;
      if (op == kConst) continue;
      if (op == kParam) {
        if (in[2] < 0 || in[2] >= npar) return VXM_ERR_INVALID;
        continue;
      }
      if (op == kUnk) {
        if (in[2] < 0 || in[2] >= m) return VXM_ERR_INVALID;
        continue;
      }
      if (in[2] < 0 || in[2] >= nreg || in[3] < 0 || in[3] >= nreg) return VXM_ERR_INVALID;
      if ((op == kSelect || op == kLimit) && (in[4] < 0 || in[4] >= nreg))
        return VXM_ERR_INVALID;
    }
    const int* out = pr + kHead + nins * kIns;
    for (int i = 0; i < 2 * m; ++i)
      if (out[i] < 0 || out[i] >= nreg) return VXM_ERR_INVALID;
    if (!generated) continue;
    int g = -1;
    for (size_t i = 0; i < gen.size(); ++i)
      if (gen[i].code_off == off) g = (int)i;
    if (g < 0) {
      g = (int)gen.size();
      VAGenProgram p;
      p.code_off = (int)off;
      p.m = m;
      p.npar = npar;
      gen.push_back(p);
    }
    gen_of[d] = g;
  }
  // device lists of the generated programs, in device order
  std::vector<int> devs;
  for (size_t g = 0; g < gen.size(); ++g) {
    gen[g].dev_off = (int)devs.size();
    for (int d = 0; d < t->D; ++d)
      if (gen_of[d] == (int)g) devs.push_back(d);
    gen[g].n_devs = (int)devs.size() - gen[g].dev_off;
  }
  for (int d = 0; d < t->D; ++d)
    if (kind_is_va(t->kind[d])) t->nterm[d] = code[t->wave[d] + 5];
  t->code = std::move(code);
  t->gen = std::move(gen);
  t->va_devs = std::move(devs);
  t->code_checked = true;
  return VXM_OK;
}

}  // namespace vxm
