// extern "C" entry points (declared in include/voltax_mi355x.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/voltax_mi355x.h"
#include "kernels.hpp"
#include "splu.hpp"
#include "status.hpp"
#include "topology.hpp"

struct vxm_topology {
  vxm::Topology* t;
};

struct vxm_batch {
  int n_inst = 0, n_items = 0;
  std::vector<int> band;  // per topology: band numbering (VXM_BAND_LU)
  int begin[vxm::kDenseClasses] = {}, count[vxm::kDenseClasses] = {};
  int count_band[vxm::kDenseClasses] = {};  // leading items of a class with a band numbering
  int lds[vxm::kDenseClasses] = {};
  void* dev = nullptr;  // one allocation: pats | items | ipoff | izoff | blobs
  const vxm::PatternDev* pats = nullptr;
  const vxm::WaveItem* items = nullptr;
  const int64_t* ipoff = nullptr;
  const int64_t* izoff = nullptr;
  // derived parameter rows of a run (parameter-only factors of the device
  // models, computed once per run); n_par = extent of the parameter buffer
  double* dparams = nullptr;
  int64_t n_par = 0;
  // one stream per tile class so the classes run concurrently; events join
  // them back to the caller's stream and time each class launch
  // (created by vxm_batch_create; vxm_batch_run only enqueues on them, under
  // `mu`, so runs from several host threads on one handle serialize instead of
  // interleaving their fork / join events)
  hipStream_t streams[vxm::kDenseClasses] = {};
  hipEvent_t ev_fork = nullptr;
  hipEvent_t ev_beg[vxm::kDenseClasses] = {}, ev_end[vxm::kDenseClasses] = {};
  std::mutex mu;
  bool timed = false;
  ~vxm_batch() {
    for (int c = 0; c < vxm::kDenseClasses; ++c) {
      if (streams[c]) (void)hipStreamDestroy(streams[c]);
      if (ev_beg[c]) (void)hipEventDestroy(ev_beg[c]);
      if (ev_end[c]) (void)hipEventDestroy(ev_end[c]);
    }
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (dev) (void)hipFree(dev);
    if (dparams) (void)hipFree(dparams);
  }
};

struct vxm_splu {
  const vxm::Topology* topo = nullptr;
  vxm::SPlanHost h;
  vxm::SPlanDev d{};
  void* dev = nullptr;
  std::mutex mu;
  ~vxm_splu() {
    if (dev) (void)hipFree(dev);
  }
};

namespace {
int upload_splu(vxm_splu* p) {
  std::lock_guard<std::mutex> lock(p->mu);
  if (p->dev) return VXM_OK;
  const vxm::SPlanHost& h = p->h;
  auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
  const std::vector<int>* ints[] = {&h.idx, &h.cmap, &h.children, &h.aslot, &h.apos,
                                    &h.level_sn};
  size_t total = al(sizeof(vxm::SNode) * h.sn.size());
  for (auto* v : ints) total += al(sizeof(int) * (v->size() + 1));
  std::vector<char> host(total, 0);
  size_t o = 0;
  std::memcpy(host.data(), h.sn.data(), sizeof(vxm::SNode) * h.sn.size());
  o += al(sizeof(vxm::SNode) * h.sn.size());
  std::vector<size_t> offs;
  for (auto* v : ints) {
    offs.push_back(o);
    if (!v->empty()) std::memcpy(host.data() + o, v->data(), sizeof(int) * v->size());
    o += al(sizeof(int) * (v->size() + 1));
  }
  void* dev = nullptr;
  if (hipMalloc(&dev, total) != hipSuccess) return VXM_ERR_HIP;
  if (hipMemcpy(dev, host.data(), total, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(dev);
    return VXM_ERR_HIP;
  }
  char* base = static_cast<char*>(dev);
  vxm::SPlanDev& d = p->d;
  d.S = h.S;
  d.nnz = h.nnz;
  d.ns = (int)h.sn.size();
  d.piv_total = h.piv_total;
  d.r_total = h.r_total;
  d.front_total = h.front_total;
  d.sn = reinterpret_cast<const vxm::SNode*>(base);
  const int** ptrs[] = {&d.idx, &d.cmap, &d.children, &d.aslot, &d.apos, &d.level_sn};
  for (size_t i = 0; i < offs.size(); ++i) *ptrs[i] = reinterpret_cast<const int*>(base + offs[i]);
  p->dev = dev;
  return VXM_OK;
}

inline hipStream_t S(vxm_stream s) { return reinterpret_cast<hipStream_t>(s); }
vxm::NewtonOpts to_opts(const vxm_newton_options* o) {
  vxm::NewtonOpts n;
  n.max_steps = o->max_steps;
  n.rtol = o->rtol;
  n.atol = o->atol;
  n.max_dv = o->max_dv;
  n.gmin = o->gmin;
  n.gmin_steps = o->gmin_steps;
  n.gmin_start = o->gmin_start;
  n.direct_steps = o->direct_steps;
  n.band_max_mult = o->band_max_mult > 0.0 ? o->band_max_mult : 100.0;
  return n;
}
}  // namespace

extern "C" {

const char* vxm_status_string(int status) {
  switch (status) {
    case VXM_OK: return "ok";
    case VXM_ERR_INVALID: return "invalid argument";
    case VXM_ERR_HIP: return "HIP runtime error";
    case VXM_ERR_MISMATCH: return "plan / topology mismatch";
    case VXM_ERR_UNSUPPORTED: return "unsupported";
    case VXM_ERR_TOO_LARGE: return "batch too large for the fixed no-allocation scratch";
    case VXM_ERR_HOST_MEMORY: return "host memory exhausted";
  }
  return "unknown status";
}

int vxm_topology_create(int32_t size, int32_t n_nodes, int32_t n_devices, int32_t n_params,
                        const int32_t* dev_kind, const int32_t* dev_wave,
                        const int32_t* dev_unknowns, const int32_t* dev_param_off,
                        vxm_topology** out) {
  return vxm_topology_create_wide(size, n_nodes, n_devices, n_params, 4, dev_kind, dev_wave,
                                  dev_unknowns, dev_param_off, out);
}

int vxm_topology_create_wide(int32_t size, int32_t n_nodes, int32_t n_devices,
                             int32_t n_params, int32_t width, const int32_t* dev_kind,
                             const int32_t* dev_wave, const int32_t* dev_unknowns,
                             const int32_t* dev_param_off, vxm_topology** out) {
  VXM_TRY
  if (!out) return VXM_ERR_INVALID;
  *out = nullptr;
  vxm::Topology* t = nullptr;
  int st = vxm::build_topology(size, n_nodes, n_devices, n_params, width, dev_kind, dev_wave,
                               dev_unknowns, dev_param_off, &t);
  if (st != VXM_OK) return st;
  auto* h = new (std::nothrow) vxm_topology{t};
  if (!h) {
    delete t;
    return VXM_ERR_INVALID;
  }
  *out = h;
  return VXM_OK;
  VXM_CATCH
}

int vxm_topology_set_code(vxm_topology* topo, int32_t n_words, const int32_t* words) {
  VXM_TRY
  if (!topo) return VXM_ERR_INVALID;
  return vxm::set_va_code(topo->t, words, n_words);
  VXM_CATCH
}

int vxm_topology_destroy(vxm_topology* topo) {
  VXM_TRY
  if (!topo) return VXM_ERR_INVALID;
  delete topo->t;
  delete topo;
  return VXM_OK;
  VXM_CATCH
}

int vxm_topology_sizes(const vxm_topology* topo, int64_t* sizes) {
  VXM_TRY
  if (!topo || !sizes) return VXM_ERR_INVALID;
  const vxm::Topology& t = *topo->t;
  sizes[0] = t.S;
  sizes[1] = t.nv;
  sizes[2] = t.D;
  sizes[3] = t.P;
  sizes[4] = t.nnz;
  sizes[5] = (int64_t)t.Cres + t.Cjac;
  return VXM_OK;
  VXM_CATCH
}

int vxm_topology_pattern(const vxm_topology* topo, int32_t* indptr, int32_t* indices) {
  VXM_TRY
  if (!topo || !indptr || !indices) return VXM_ERR_INVALID;
  const vxm::Topology& t = *topo->t;
  std::memcpy(indptr, t.indptr.data(), sizeof(int32_t) * (t.S + 1));
  std::memcpy(indices, t.indices.data(), sizeof(int32_t) * t.nnz);
  return VXM_OK;
  VXM_CATCH
}

int vxm_stamp(const vxm_topology* topo, int32_t batch, const double* params, const double* z,
              double t, double cf, double cq, double gmin, const double* hist, double* data,
              double* res, double* work, vxm_stream stream) {
  VXM_TRY
  if (!topo || batch < 0 || !z || !work || (!data && !res)) return VXM_ERR_INVALID;
  if (topo->t->P > 0 && !params) return VXM_ERR_INVALID;
  if (batch == 0) return VXM_OK;  // an empty batch: nothing to launch
  if (int st = vxm::upload_topology(topo->t)) return st;
  return vxm::launch_stamp(topo->t->dev, batch, params, z, t, cf, cq, gmin, hist, data, res,
                           work, S(stream));
  VXM_CATCH
}

int vxm_stamp_dev(const vxm_topology* topo, int32_t batch, const double* params,
                  const double* z, const double* scal, const double* hist, double* data,
                  double* res, double* work, vxm_stream stream) {
  VXM_TRY
  if (!topo || batch < 0 || !z || !scal || !work || (!data && !res)) return VXM_ERR_INVALID;
  if (topo->t->P > 0 && !params) return VXM_ERR_INVALID;
  if (batch == 0) return VXM_OK;  // an empty batch: nothing to launch
  if (int st = vxm::upload_topology(topo->t)) return st;
  return vxm::launch_stamp(topo->t->dev, batch, params, z, 0.0, 0.0, 0.0, 0.0, hist, data, res,
                           work, S(stream), scal);
  VXM_CATCH
}

int vxm_spmv(const vxm_topology* topo, int32_t batch, const double* data, const double* x,
             double* y, vxm_stream stream) {
  VXM_TRY
  if (!topo || batch < 0 || !data || !x || !y) return VXM_ERR_INVALID;
  if (batch == 0) return VXM_OK;  // an empty batch: nothing to launch
  if (int st = vxm::upload_topology(topo->t)) return st;
  return vxm::launch_spmv(topo->t->dev, batch, data, x, y, S(stream));
  VXM_CATCH
}

int vxm_param_vjp(const vxm_topology* topo, int32_t batch, const double* params,
                  const double* z, double t, double cf, double cq, const double* lam,
                  double scale, double* grad, vxm_stream stream) {
  VXM_TRY
  if (!topo || batch < 0 || !z || !lam || !grad) return VXM_ERR_INVALID;
  if (topo->t->P > 0 && !params) return VXM_ERR_INVALID;
  if (batch == 0) return VXM_OK;  // an empty batch: nothing to launch
  if (int st = vxm::upload_topology(topo->t)) return st;
  return vxm::launch_param_vjp(topo->t->dev, batch, params, z, t, cf, cq, lam, scale, grad,
                               S(stream));
  VXM_CATCH
}

int vxm_time_vjp(const vxm_topology* topo, int32_t batch, const double* params, double t,
                 double cf, const double* lam, double scale, double* out, vxm_stream stream) {
  VXM_TRY
  if (!topo || batch < 0 || !lam || !out) return VXM_ERR_INVALID;
  if (topo->t->P > 0 && !params) return VXM_ERR_INVALID;
  if (batch == 0) return VXM_OK;  // an empty batch: nothing to launch
  if (int st = vxm::upload_topology(topo->t)) return st;
  return vxm::launch_time_vjp(topo->t->dev, batch, params, t, cf, lam, scale, out, S(stream));
  VXM_CATCH
}

int vxm_jt_vec(const vxm_topology* topo, int32_t batch, const double* params, const double* z,
               double t, double cf, double cq, double gmin, const double* lam, double* out,
               double* work, vxm_stream stream) {
  VXM_TRY
  if (!topo || batch < 0 || !z || !lam || !out || !work) return VXM_ERR_INVALID;
  if (topo->t->P > 0 && !params) return VXM_ERR_INVALID;
  if (batch == 0) return VXM_OK;  // an empty batch: nothing to launch
  if (int st = vxm::upload_topology(topo->t)) return st;
  return vxm::launch_jt_vec(topo->t->dev, batch, params, z, t, cf, cq, gmin, lam, out, work,
                            S(stream));
  VXM_CATCH
}

namespace {
// Reverse Cuthill-McKee order of the unknowns outside `skip` (adjacency as
// bit masks; S <= 64): components in turn, each from a pseudo-peripheral
// vertex of least degree, neighbours visited in ascending degree.
std::vector<int> rcm_order(const std::vector<unsigned long long>& adj, unsigned long long skip) {
  const int S = (int)adj.size();
  auto deg = [&](int u) { return __builtin_popcountll(adj[u] & ~skip & ~(1ull << u)); };
  unsigned long long seen = skip;
  std::vector<int> order;
  auto bfs = [&](int start, unsigned long long taken, std::vector<int>* out) {
    // returns the vertices of the last level (and appends the visit order)
    std::vector<int> level{start}, last;
    taken |= 1ull << start;
    if (out) out->push_back(start);
    while (!level.empty()) {
      last = level;
      std::vector<int> next;
      for (int u : level) {
        std::vector<int> nb;
        for (int v = 0; v < S; ++v)
          if (v != u && ((adj[u] >> v) & 1) && !((taken >> v) & 1)) nb.push_back(v);
        std::stable_sort(nb.begin(), nb.end(), [&](int a, int b) { return deg(a) < deg(b); });
        for (int v : nb) {
          taken |= 1ull << v;
          next.push_back(v);
          if (out) out->push_back(v);
        }
      }
      level = next;
    }
    return last;
  };
  while (true) {
    int start = -1;
    for (int u = 0; u < S; ++u)
      if (!((seen >> u) & 1) && (start < 0 || deg(u) < deg(start))) start = u;
    if (start < 0) break;
    for (int it = 0; it < 2; ++it) {  // pseudo-peripheral start
      std::vector<int> last = bfs(start, seen, nullptr);
      int best = last[0];
      for (int v : last)
        if (deg(v) < deg(best)) best = v;
      if (best == start) break;
      start = best;
    }
    std::vector<int> comp;
    bfs(start, seen, &comp);
    for (int v : comp) seen |= 1ull << v;
    order.insert(order.end(), comp.begin(), comp.end());
  }
  std::reverse(order.begin(), order.end());
  return order;
}

// Numbering of a pattern for the static-pivot band LU (kBandWidth,
// kBandBorder in kernels.hpp): the nodes and branch currents of voltage
// sources (a branch row has a zero diagonal) and, while the band is too wide,
// the remaining node of highest degree form the border at the last
// positions; the other unknowns follow reverse Cuthill-McKee.  pos[u]: the
// position of unknown u.  False if the pattern does not fit.
bool band_numbering(const vxm::Topology& t, int smax, std::vector<int>& pos) {
  const int S = t.S;
  if (S > smax || smax > 64) return false;
  std::vector<unsigned long long> adj(S, 0);
  unsigned long long border = 0;
  for (int d = 0; d < t.D; ++d) {
    const int m = vxm::kind_m(t.kind[d]);
    const int* u = t.unk.data() + (size_t)t.uw * d;
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < m; ++j)
        if (u[i] >= 0 && u[j] >= 0) adj[u[i]] |= 1ull << u[j];
    if (t.kind[d] == vxm::KIND_V)
      for (int i = 0; i < m; ++i)
        if (u[i] >= 0) border |= 1ull << u[i];
  }
  const int KB = smax - vxm::kBandBorder;  // first border-zone position
  while (__builtin_popcountll(border) <= vxm::kBandBorder) {
    const std::vector<int> order = rcm_order(adj, border);
    pos.assign(S, -1);
    int p = 0;
    for (int u : order) pos[u] = p++;
    std::vector<int> bl;
    for (int u = 0; u < S; ++u)
      if ((border >> u) & 1) bl.push_back(u);
    std::stable_sort(bl.begin(), bl.end(), [&](int a, int b) {  // nodes before currents
      return (a < t.nv) > (b < t.nv);
    });
    int q = smax - (int)bl.size();
    for (int u : bl) pos[u] = q++;
    bool ok = true;
    for (int u = 0; u < S && ok; ++u)
      for (int v = 0; v < S && ok; ++v)
        if (((adj[u] >> v) & 1) && pos[u] < KB && pos[v] > pos[u] + vxm::kBandWidth &&
            pos[v] < KB)
          ok = false;
    if (ok) return true;
    int hub = -1, hd = -1;  // widen the border by the busiest remaining node
    for (int u = 0; u < t.nv; ++u) {
      if ((border >> u) & 1) continue;
      const int dg = __builtin_popcountll(adj[u] & ~border);
      if (dg > hd) hub = u, hd = dg;
    }
    if (hub < 0) return false;
    border |= 1ull << hub;
  }
  return false;
}

// int16 tables of one pattern for the fused kernel (layout in kernels.hpp)
bool pattern_blob(const vxm::Topology& t, int smax, int lanes, std::vector<short>& b,
                  vxm::PatternDev& pd) {
  if (t.Cjac > 32767 || t.Cres > 32767 || t.P > 32767) return false;
  // positions of the unknowns: a band numbering when the pattern has one
  std::vector<int> pos;
  pd.band = band_numbering(t, smax, pos) ? 1 : 0;
  if (!pd.band) {
    pos.resize(t.S);
    for (int u = 0; u < t.S; ++u) pos[u] = u;
  }
  pd.sk = pd.band ? smax : t.S;
  pd.nodemask = pd.realmask = 0;
  for (int u = 0; u < t.S; ++u) {
    pd.realmask |= 1ull << pos[u];
    if (u < t.nv) pd.nodemask |= 1ull << pos[u];
  }
  auto at = [&](int u) { return u >= 0 ? pos[u] : -1; };
  // device rows grouped by kind (EKV first, the expensive one), each kind
  // padded to a whole number of passes so no pass mixes kinds, unless the
  // padding would cost extra passes (tiny circuits): then kinds stay grouped
  // but unpadded
  std::vector<int> rows;
  const int kinds[] = {vxm::KIND_EKV, vxm::KIND_C, vxm::KIND_R, vxm::KIND_V, vxm::KIND_I};
  int aligned = 0;
  for (int k : kinds) {
    int n = 0;
    for (int d = 0; d < t.D; ++d) n += t.kind[d] == k;
    aligned += (n + lanes - 1) / lanes;
  }
  const bool pad = aligned <= (t.D + lanes - 1) / lanes;
  for (int k : kinds) {
    size_t before = rows.size();
    for (int d = 0; d < t.D; ++d)
      if (t.kind[d] == k) rows.push_back(d);
    while (pad && (rows.size() - before) % lanes) rows.push_back(-1);
  }
  pd.dev_slots = (int)rows.size();
  b.assign(rows.size() * 10, 0);
  for (size_t i = 0; i < rows.size(); ++i) {
    short* r = b.data() + 10 * i;
    const int d = rows[i];
    if (d < 0) {
      r[0] = -1;
      continue;
    }
    r[0] = (short)t.kind[d];
    r[1] = (short)t.wave[d];
    for (int j = 0; j < 4; ++j) r[2 + j] = (short)at(t.unk[t.uw * d + j]);
    r[6] = (short)t.poff[d];
    r[7] = (short)t.roff[d];
    r[8] = (short)t.joff[d];
  }
  // every device adds its residual and Jacobian entries straight into the
  // instance's residual vector and column-major sk x sk matrix in LDS (at
  // the positions of its unknowns); entries on a ground row or column are
  // dropped
  const int sk = pd.sk;
  pd.o_jpos = (int)b.size();
  b.resize(b.size() + t.Cjac, (short)-1);
  pd.o_rpos = (int)b.size();
  b.resize(b.size() + t.Cres, (short)-1);
  for (int d = 0; d < t.D; ++d) {
    const int m = vxm::kind_m(t.kind[d]);
    const int* u = t.unk.data() + (size_t)t.uw * d;
    for (int i = 0; i < m; ++i) {
      if (u[i] < 0) continue;
      b[pd.o_rpos + t.roff[d] + i] = (short)pos[u[i]];
      for (int j = 0; j < m; ++j)
        if (u[j] >= 0)
          b[pd.o_jpos + t.joff[d] + i * m + j] = (short)(pos[u[j]] * sk + pos[u[i]]);
    }
  }
  pd.o_q = (int)b.size();
  b.resize(b.size() + smax, (short)-1);
  for (int u = 0; u < t.S; ++u) b[pd.o_q + pos[u]] = (short)u;
  if (b.size() & 1) b.push_back(0);
  pd.blob_words = (int)(b.size() / 2);
  return true;
}
}  // namespace

int vxm_batch_create(int32_t n_topos, vxm_topology* const* topos, int32_t n_inst,
                     const int32_t* inst_topo, const int64_t* inst_param_off,
                     vxm_batch** out) {
  VXM_TRY
  if (!out || n_topos <= 0 || !topos || n_inst <= 0 || !inst_topo || !inst_param_off)
    return VXM_ERR_INVALID;
  *out = nullptr;
  std::vector<vxm::PatternDev> pats(n_topos);
  std::vector<std::vector<short>> blobs(n_topos);
  int lds[vxm::kDenseClasses] = {};
  for (int p = 0; p < n_topos; ++p) {
    if (!topos[p]) return VXM_ERR_INVALID;
    for (int k : topos[p]->t->kind)
      if (vxm::kind_is_va(k)) return VXM_ERR_UNSUPPORTED;  // Verilog-A: stamped (sparse) path
    {  // the derived rows of a run are written in place of these parameter rows
      const vxm::Topology& tp = *topos[p]->t;
      for (int d = 0; d < tp.D; ++d) {
        const int need = tp.kind[d] == vxm::KIND_EKV ? (int)vxm::EKV_NPARAM
                         : (tp.kind[d] == vxm::KIND_R || tp.kind[d] == vxm::KIND_C) ? 1 : 0;
        if (tp.poff[d] + need > tp.P) return VXM_ERR_INVALID;
      }
    }
    if (int st = vxm::upload_topology(topos[p]->t)) return st;
    const vxm::Topology& t = *topos[p]->t;
    int c = vxm::dense_class(t.S);
    if (c < 0) return VXM_ERR_UNSUPPORTED;
    const int smax = vxm::kDenseSmax[c], lanes = vxm::kDenseLanes[c];
    pats[p].t = t.dev;
    pats[p].smax = smax;
    pats[p].lanes = lanes;
    pats[p].cls = c;
    if (!pattern_blob(t, smax, lanes, blobs[p], pats[p])) return VXM_ERR_UNSUPPORTED;
    // zl[smax] | res[sk] | matrix[smax columns of sk] (columns sk.. stay zero) | pivots
    const int sk = pats[p].sk;
    pats[p].seg_doubles = smax + sk + smax * sk + (smax + 1) / 2 + 1;
    int bytes = (64 / lanes) * pats[p].seg_doubles * (int)sizeof(double) +
                ((pats[p].blob_words * 4 + 7) / 8) * 8;
    if (bytes > 160 * 1024) return VXM_ERR_UNSUPPORTED;
    lds[c] = std::max(lds[c], bytes);
  }
  std::vector<int64_t> izoff(n_inst);
  int64_t acc = 0, n_par = 0;
  std::vector<std::vector<int>> by_pat(n_topos);
  for (int i = 0; i < n_inst; ++i) {
    int p = inst_topo[i];
    if (p < 0 || p >= n_topos || inst_param_off[i] < 0) return VXM_ERR_INVALID;
    n_par = std::max<int64_t>(n_par, inst_param_off[i] + topos[p]->t->P);
    izoff[i] = acc;
    acc += topos[p]->t->S;
    by_pat[p].push_back(i);
  }
  // wave items grouped by tile class; inside a class, patterns with a band
  // numbering first, then heaviest patterns first
  std::vector<int> order(n_topos);
  for (int p = 0; p < n_topos; ++p) order[p] = p;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    if (pats[a].cls != pats[b].cls) return pats[a].cls < pats[b].cls;
    if (pats[a].band != pats[b].band) return pats[a].band > pats[b].band;
    return topos[a]->t->S > topos[b]->t->S;
  });
  std::vector<vxm::WaveItem> items;
  for (int p : order) {
    const int per = 64 / pats[p].lanes;  // instances of this pattern per wave
    const auto& v = by_pat[p];
    for (size_t i = 0; i < v.size(); i += per) {
      vxm::WaveItem w{};
      w.pattern = p;
      w.ninst = (int)std::min(v.size() - i, (size_t)per);
      for (int k = 0; k < w.ninst; ++k) w.inst[k] = v[i + k];
      items.push_back(w);
    }
  }
  auto* b = new (std::nothrow) vxm_batch();
  if (!b) return VXM_ERR_INVALID;
  for (size_t i = 0; i < items.size(); ++i) {
    int c = pats[items[i].pattern].cls;
    if (b->count[c] == 0) b->begin[c] = (int)i;
    b->count[c]++;
    b->count_band[c] += pats[items[i].pattern].band;
  }
  for (int c = 0; c < vxm::kDenseClasses; ++c) b->lds[c] = lds[c];
  for (int p = 0; p < n_topos; ++p) b->band.push_back(pats[p].band);
  auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
  size_t sz_p = sizeof(vxm::PatternDev) * pats.size();
  size_t sz_i = sizeof(vxm::WaveItem) * items.size();
  size_t sz_o = sizeof(int64_t) * n_inst;
  std::vector<size_t> boff(n_topos);
  size_t sz_b = 0;
  for (int p = 0; p < n_topos; ++p) {
    boff[p] = sz_b;
    sz_b += al(blobs[p].size() * sizeof(short));
  }
  size_t total = al(sz_p) + al(sz_i) + 2 * al(sz_o) + sz_b;
  if (hipMalloc(&b->dev, total) != hipSuccess) {
    b->dev = nullptr;
    delete b;
    return VXM_ERR_HIP;
  }
  char* base = static_cast<char*>(b->dev);
  char* bbase = base + al(sz_p) + al(sz_i) + 2 * al(sz_o);
  for (int p = 0; p < n_topos; ++p)
    pats[p].blob = reinterpret_cast<const short*>(bbase + boff[p]);
  std::vector<char> host(total, 0);
  std::memcpy(host.data(), pats.data(), sz_p);
  std::memcpy(host.data() + al(sz_p), items.data(), sz_i);
  std::memcpy(host.data() + al(sz_p) + al(sz_i), inst_param_off, sz_o);
  std::memcpy(host.data() + al(sz_p) + al(sz_i) + al(sz_o), izoff.data(), sz_o);
  for (int p = 0; p < n_topos; ++p)
    std::memcpy(host.data() + (bbase - base) + boff[p], blobs[p].data(),
                blobs[p].size() * sizeof(short));
  if (hipMemcpy(b->dev, host.data(), total, hipMemcpyHostToDevice) != hipSuccess) {
    delete b;
    return VXM_ERR_HIP;
  }
  b->pats = reinterpret_cast<const vxm::PatternDev*>(base);
  b->items = reinterpret_cast<const vxm::WaveItem*>(base + al(sz_p));
  b->ipoff = reinterpret_cast<const int64_t*>(base + al(sz_p) + al(sz_i));
  b->izoff = reinterpret_cast<const int64_t*>(base + al(sz_p) + al(sz_i) + al(sz_o));
  b->n_inst = n_inst;
  b->n_items = (int)items.size();
  b->n_par = n_par;
  if (n_par > 0 && hipMalloc(&b->dparams, sizeof(double) * n_par) != hipSuccess) {
    b->dparams = nullptr;
    delete b;
    return VXM_ERR_HIP;
  }
  if (hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess) {
    b->ev_fork = nullptr;
    delete b;
    return VXM_ERR_HIP;
  }
  for (int c = 0; c < vxm::kDenseClasses; ++c) {
    if (!b->count[c]) continue;
    if (hipStreamCreateWithFlags(&b->streams[c], hipStreamNonBlocking) != hipSuccess) {
      b->streams[c] = nullptr;
      delete b;  // frees what was created so far
      return VXM_ERR_HIP;
    }
    if (hipEventCreate(&b->ev_beg[c]) != hipSuccess) {
      b->ev_beg[c] = nullptr;
      delete b;
      return VXM_ERR_HIP;
    }
    if (hipEventCreate(&b->ev_end[c]) != hipSuccess) {
      b->ev_end[c] = nullptr;
      delete b;
      return VXM_ERR_HIP;
    }
  }
  *out = b;
  return VXM_OK;
  VXM_CATCH
}

int vxm_batch_destroy(vxm_batch* batch) {
  VXM_TRY
  if (!batch) return VXM_ERR_INVALID;
  delete batch;
  return VXM_OK;
  VXM_CATCH
}

int vxm_batch_run(const vxm_batch* batch, const double* params, const double* z0,
                  const double* ts, int32_t T, int32_t method, int32_t do_dc, double t_dc,
                  const vxm_newton_options* opts, int32_t precision, double* out_z,
                  int8_t* out_conv, int32_t* out_iters, vxm_stream stream) {
  VXM_TRY
  if (!batch || !z0 || !opts || !out_z || !out_conv || !out_iters || T < 1) return VXM_ERR_INVALID;
  if (T > 1 && !ts) return VXM_ERR_INVALID;
  const int32_t base = method & ~(VXM_PREDICT | VXM_REUSE | VXM_BAND_LU);
  if (base != 0 && base != 1) return VXM_ERR_INVALID;
  if (precision != 0 && precision != 1) return VXM_ERR_INVALID;
  if (method & VXM_REUSE) {  // FP64 factors kept across the iterations of a step
    if (precision != 0) return VXM_ERR_INVALID;
    precision = 2;
  }
  if (opts->max_steps < 0 || opts->gmin_steps < 0 || opts->direct_steps < 0 ||
      !(opts->band_max_mult >= 0.0))
    return VXM_ERR_INVALID;
  auto* b = const_cast<vxm_batch*>(batch);  // only the mutex, the events and `timed` change
  std::lock_guard<std::mutex> lock(b->mu);
  if (b->n_par > 0 && !params) return VXM_ERR_INVALID;
  hipStream_t caller = S(stream);
  // the derived rows are rewritten only after the previous run's launches
  // (which may be on other streams) are done with them
  for (int c = 0; c < vxm::kDenseClasses; ++c)
    if (b->count[c] && hipStreamWaitEvent(caller, b->ev_end[c], 0) != hipSuccess)
      return VXM_ERR_HIP;
  if (int st = vxm::launch_derive_params(b->pats, b->items, b->n_items, b->ipoff, b->n_par,
                                         params, b->dparams, caller))
    return st;
  if (hipEventRecord(b->ev_fork, caller) != hipSuccess) return VXM_ERR_HIP;
  const vxm::NewtonOpts o = to_opts(opts);
  for (int c = vxm::kDenseClasses - 1; c >= 0; --c) {  // largest tiles first
    if (!b->count[c]) continue;
    hipStream_t sc = b->streams[c];
    if (hipStreamWaitEvent(sc, b->ev_fork, 0) != hipSuccess) return VXM_ERR_HIP;
    if (hipEventRecord(b->ev_beg[c], sc) != hipSuccess) return VXM_ERR_HIP;
    int cnt[vxm::kDenseClasses] = {}, beg[vxm::kDenseClasses] = {};
    int cnb[vxm::kDenseClasses] = {};
    cnt[c] = b->count[c];
    beg[c] = b->begin[c];
    cnb[c] = b->count_band[c];
    int st = vxm::launch_dense_run(b->pats, b->items, beg, cnt, cnb, b->lds, b->ipoff, b->izoff,
                                   b->dparams, z0, ts, T, method, do_dc, t_dc, o, precision, out_z,
                                   out_conv, out_iters, sc);
    if (st) return st;
    if (hipEventRecord(b->ev_end[c], sc) != hipSuccess) return VXM_ERR_HIP;
    if (hipStreamWaitEvent(caller, b->ev_end[c], 0) != hipSuccess) return VXM_ERR_HIP;
  }
  b->timed = true;
  return VXM_OK;
  VXM_CATCH
}

int vxm_batch_timings(const vxm_batch* batch, float* ms) {
  VXM_TRY
  if (!batch || !ms) return VXM_ERR_INVALID;
  std::lock_guard<std::mutex> lock(const_cast<vxm_batch*>(batch)->mu);
  for (int c = 0; c < vxm::kDenseClasses; ++c) {
    ms[c] = 0.0f;
    if (!batch->count[c] || !batch->timed) continue;
    if (hipEventSynchronize(batch->ev_end[c]) != hipSuccess) return VXM_ERR_HIP;
    if (hipEventElapsedTime(&ms[c], batch->ev_beg[c], batch->ev_end[c]) != hipSuccess)
      return VXM_ERR_HIP;
  }
  return VXM_OK;
  VXM_CATCH
}

int vxm_topology_band(const vxm_topology* topo, int32_t smax, int32_t* pos) {
  VXM_TRY
  if (!topo || !pos || smax < 1 || smax > 64) return VXM_ERR_INVALID;
  std::vector<int> p;
  if (!band_numbering(*topo->t, smax, p)) return VXM_ERR_UNSUPPORTED;
  for (size_t u = 0; u < p.size(); ++u) pos[u] = p[u];
  return VXM_OK;
  VXM_CATCH
}

int vxm_batch_band(const vxm_batch* batch, int32_t* band) {
  VXM_TRY
  if (!batch || !band) return VXM_ERR_INVALID;
  for (size_t p = 0; p < batch->band.size(); ++p) band[p] = batch->band[p];
  return VXM_OK;
  VXM_CATCH
}

int vxm_splu_create_ex(const vxm_topology* topo, int32_t leaf, int32_t flags, vxm_splu** out) {
  VXM_TRY
  if (!topo || !out || leaf < 1 || (flags & ~VXM_SPLU_COMPACT)) return VXM_ERR_INVALID;
  *out = nullptr;
  auto* p = new (std::nothrow) vxm_splu();
  if (!p) return VXM_ERR_INVALID;
  p->topo = topo->t;
  int st = vxm::build_splu_plan(*topo->t, leaf, &p->h, (flags & VXM_SPLU_COMPACT) != 0);
  if (st != VXM_OK) {
    delete p;
    return st;
  }
  *out = p;
  return VXM_OK;
  VXM_CATCH
}

int vxm_splu_create(const vxm_topology* topo, int32_t leaf, vxm_splu** out) {
  return vxm_splu_create_ex(topo, leaf, 0, out);
}

int vxm_splu_destroy(vxm_splu* plan) {
  VXM_TRY
  if (!plan) return VXM_ERR_INVALID;
  delete plan;
  return VXM_OK;
  VXM_CATCH
}

int vxm_splu_sizes(const vxm_splu* plan, int64_t* sizes) {
  VXM_TRY
  if (!plan || !sizes) return VXM_ERR_INVALID;
  const vxm::SPlanHost& h = plan->h;
  sizes[0] = h.front_total;
  sizes[1] = h.piv_total;
  sizes[2] = h.r_total;
  sizes[3] = (int64_t)h.sn.size();
  sizes[4] = (int64_t)h.level_ptr.size() - 1;
  sizes[5] = h.max_front;
  sizes[6] = h.max_k;
  sizes[7] = (int64_t)h.r_total + h.S;
  return VXM_OK;
  VXM_CATCH
}

int vxm_splu_stats(const vxm_splu* plan, double* stats) {
  VXM_TRY
  if (!plan || !stats) return VXM_ERR_INVALID;
  stats[0] = plan->h.lu_flops;
  stats[1] = plan->h.factor_entries;
  return VXM_OK;
  VXM_CATCH
}

int vxm_splu_structure(const vxm_splu* plan, int32_t* sn_of, int32_t* parent, int32_t* level) {
  VXM_TRY
  if (!plan || !sn_of || !parent || !level) return VXM_ERR_INVALID;
  const vxm::SPlanHost& h = plan->h;
  for (size_t i = 0; i < h.sn.size(); ++i) {
    const vxm::SNode& s = h.sn[i];
    parent[i] = s.parent;
    level[i] = s.level;
    for (int j = 0; j < s.k; ++j) sn_of[h.idx[s.idx_off + j]] = (int32_t)i;
  }
  return VXM_OK;
  VXM_CATCH
}

int vxm_splu_export(const vxm_splu* plan, int64_t* counts, int64_t* sn, int32_t* sn2,
                    int32_t* idx, int32_t* cmap, int32_t* children, int32_t* aslot,
                    int32_t* apos, int32_t* level_ptr, int32_t* level_sn) {
  VXM_TRY
  if (!plan || !counts) return VXM_ERR_INVALID;
  const vxm::SPlanHost& h = plan->h;
  counts[0] = (int64_t)h.sn.size();
  counts[1] = (int64_t)h.idx.size();
  counts[2] = (int64_t)h.cmap.size();
  counts[3] = (int64_t)h.children.size();
  counts[4] = (int64_t)h.aslot.size();
  counts[5] = (int64_t)h.level_sn.size();
  if (!sn) return VXM_OK;
  if (!sn2 || !idx || !cmap || !children || !aslot || !apos || !level_ptr || !level_sn)
    return VXM_ERR_INVALID;
  for (size_t i = 0; i < h.sn.size(); ++i) {
    const vxm::SNode& s = h.sn[i];
    const int64_t row[10] = {s.k, s.r, s.parent, s.level, s.front_off, s.piv_off,
                             s.idx_off, s.cmap_off, s.child_off, s.n_child};
    std::memcpy(sn + 10 * i, row, sizeof(row));
    sn2[3 * i] = s.aent_off;
    sn2[3 * i + 1] = s.n_aent;
    sn2[3 * i + 2] = s.uvec_off;
  }
  std::copy(h.idx.begin(), h.idx.end(), idx);
  std::copy(h.cmap.begin(), h.cmap.end(), cmap);
  std::copy(h.children.begin(), h.children.end(), children);
  std::copy(h.aslot.begin(), h.aslot.end(), aslot);
  std::copy(h.apos.begin(), h.apos.end(), apos);
  std::copy(h.level_ptr.begin(), h.level_ptr.end(), level_ptr);
  std::copy(h.level_sn.begin(), h.level_sn.end(), level_sn);
  return VXM_OK;
  VXM_CATCH
}

int vxm_dense_classes(int32_t* n_classes, int32_t* smax, int32_t* lanes) {
  VXM_TRY
  if (!n_classes) return VXM_ERR_INVALID;
  *n_classes = vxm::kDenseClasses;
  for (int c = 0; c < vxm::kDenseClasses; ++c) {
    if (smax) smax[c] = vxm::kDenseSmax[c];
    if (lanes) lanes[c] = vxm::kDenseLanes[c];
  }
  return VXM_OK;
  VXM_CATCH
}

int vxm_splu_configure(vxm_splu* plan, int32_t mid_blocked) {
  VXM_TRY
  if (!plan || (mid_blocked != 0 && mid_blocked != 1)) return VXM_ERR_INVALID;
  std::lock_guard<std::mutex> lock(plan->mu);
  plan->h.mid_blocked = mid_blocked;
  return VXM_OK;
  VXM_CATCH
}

int vxm_splu_factor(const vxm_splu* plan, int32_t batch, const double* data, void* front,
                    int32_t* piv, int32_t dtype, vxm_stream stream) {
  VXM_TRY
  if (!plan || batch < 0 || !data || !front || !piv || dtype < 0 || dtype > 2)
    return VXM_ERR_INVALID;
  if (batch == 0) return VXM_OK;  // an empty batch: nothing to launch
  auto* p = const_cast<vxm_splu*>(plan);
  if (int st = upload_splu(p)) return st;
  return vxm::splu_factor(p->h, p->d, batch, data, front, piv, dtype, S(stream));
  VXM_CATCH
}

int vxm_splu_solve(const vxm_splu* plan, int32_t batch, const void* front, const int32_t* piv,
                   double* x, double* work, int32_t transpose, int32_t dtype,
                   vxm_stream stream) {
  VXM_TRY
  if (!plan || batch < 0 || !front || !piv || !x || !work || dtype < 0 || dtype > 2)
    return VXM_ERR_INVALID;
  if (batch == 0) return VXM_OK;  // an empty batch: nothing to launch
  auto* p = const_cast<vxm_splu*>(plan);
  if (int st = upload_splu(p)) return st;
  return vxm::splu_solve(p->h, p->d, batch, front, piv, x, work, transpose != 0, dtype,
                         S(stream));
  VXM_CATCH
}

int vxm_splu_factor_solve(const vxm_splu* plan, int32_t batch, const double* data, double* front,
                          int32_t* piv, double* x, double* work, vxm_stream stream) {
  VXM_TRY
  if (!plan || batch < 0 || !data || !front || !piv || !x || !work) return VXM_ERR_INVALID;
  if (batch == 0) return VXM_OK;  // an empty batch: nothing to launch
  auto* p = const_cast<vxm_splu*>(plan);
  if (int st = upload_splu(p)) return st;
  return vxm::splu_factor_solve(p->h, p->d, batch, data, front, piv, x, work, S(stream));
  VXM_CATCH
}

int vxm_splu_solve_refined(const vxm_splu* plan, int32_t batch, const double* data,
                           const void* front, const int32_t* piv, const double* b, double* x,
                           double* work, int32_t iters, int32_t transpose, int32_t fp32,
                           vxm_stream stream) {
  VXM_TRY
  if (!plan || batch < 0 || !data || !front || !piv || !b || !x || !work || iters < 0)
    return VXM_ERR_INVALID;
  if (fp32 < 0 || fp32 > 1) return VXM_ERR_INVALID;  // real systems only
  if (transpose) return VXM_ERR_UNSUPPORTED;  // refinement of A^T solves: not needed yet
  if (batch == 0) return VXM_OK;  // an empty batch: nothing to launch
  auto* p = const_cast<vxm_splu*>(plan);
  if (int st = upload_splu(p)) return st;
  if (int st = vxm::upload_topology(const_cast<vxm::Topology*>(p->topo))) return st;
  const int64_t n = (int64_t)batch * p->h.S;
  double* r = work + (int64_t)batch * (p->h.r_total + p->h.S);
  hipStream_t st = S(stream);
  if (hipMemcpyAsync(x, b, sizeof(double) * n, hipMemcpyDeviceToDevice, st) != hipSuccess)
    return VXM_ERR_HIP;
  int rc = vxm_splu_solve(plan, batch, front, piv, x, work, 0, fp32, stream);
  for (int it = 0; it < iters && rc == 0; ++it) {
    rc = vxm::launch_residual(p->topo->dev, batch, data, x, b, r, st);
    if (rc == 0) rc = vxm_splu_solve(plan, batch, front, piv, r, work, 0, fp32, stream);
    if (rc == 0) rc = vxm::launch_axpy(n, 1.0, r, x, st);
  }
  return rc;
  VXM_CATCH
}

int vxm_ac_system(int32_t nnz, int32_t size, int32_t n_freq, const double* g, const double* c,
                  const double* freqs, const double* b, double* data, double* rhs,
                  vxm_stream stream) {
  VXM_TRY
  if (nnz <= 0 || size <= 0 || n_freq < 0 || !g || !c || !freqs || !b || !data || !rhs)
    return VXM_ERR_INVALID;
  return vxm::launch_ac_system(nnz, size, n_freq, g, c, freqs, b, data, rhs, S(stream));
  VXM_CATCH
}

int vxm_newton_update(int32_t batch, int32_t size, int32_t n_nodes, double* z,
                      const double* dz, double dz_sign, const vxm_newton_options* opts,
                      const int8_t* active, int8_t* done, double* step_norm,
                      vxm_stream stream) {
  VXM_TRY
  if (batch < 0 || size <= 0 || n_nodes < 0 || n_nodes > size || !z || !dz || !opts || !done)
    return VXM_ERR_INVALID;
  if (batch == 0) return VXM_OK;  // an empty batch: nothing to launch
  return vxm::launch_newton_update(batch, size, n_nodes, z, dz, dz_sign, to_opts(opts), active,
                                   done, step_norm, S(stream));
  VXM_CATCH
}

int vxm_newton_update_reuse(int32_t batch, int32_t size, int32_t n_nodes, double* z,
                            const double* dz, double dz_sign, const vxm_newton_options* opts,
                            const int8_t* active, int8_t* done, double* wnorm,
                            int8_t* refactor, int32_t reused, vxm_stream stream) {
  VXM_TRY
  if (batch < 0 || size <= 0 || n_nodes < 0 || n_nodes > size || !z || !dz || !opts || !done ||
      !wnorm || !refactor || (reused != 0 && reused != 1))
    return VXM_ERR_INVALID;
  if (batch == 0) return VXM_OK;  // an empty batch: nothing to launch
  return vxm::launch_newton_update_reuse(batch, size, n_nodes, z, dz, dz_sign, to_opts(opts),
                                         active, done, wnorm, refactor, reused, S(stream));
  VXM_CATCH
}

}  // extern "C"

// ----------------------------------------------------- no-allocation stamping
// The fixed scratch of vxm_stamp_noalloc: one copy per device, held until the
// process ends.  Host code never passes `scratch` itself to a kernel (on the
// host the name is only a stand-in for the symbol); it passes the device
// address that hipGetSymbolAddress returns on each device, taken once per
// device with that device current.
namespace {
static __device__ __align__(256) double scratch[VXM_NOALLOC_SCRATCH_DOUBLES];
constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
double* scratch_ptr[kMaxDevices] = {};

int scratch_on_current_device(double** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return VXM_ERR_HIP;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return VXM_ERR_HIP;
    scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = scratch_ptr[dev];
  return VXM_OK;
}
}  // namespace

extern "C" int vxm_stamp_noalloc(const vxm_topology* topo, int32_t batch, const double* params,
                                 const double* z, double t, double cf, double cq, double gmin,
                                 const double* hist, double* data, double* res,
                                 vxm_stream stream) {
  try {
    if (!topo || batch < 0 || !z || (!data && !res)) return VXM_ERR_INVALID;
    const vxm::Topology& tp = *topo->t;
    if (tp.P > 0 && !params) return VXM_ERR_INVALID;
    const int64_t need = (int64_t)batch * (tp.Cres + tp.Cjac);
    if (need > VXM_NOALLOC_SCRATCH_DOUBLES) return VXM_ERR_TOO_LARGE;
    if (batch == 0) return VXM_OK;  // an empty batch: nothing to launch
    if (!tp.dev_block) return VXM_ERR_INVALID;  // tables not resident: no upload here
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return VXM_ERR_HIP;
    if (dev != tp.dev_ordinal) return VXM_ERR_MISMATCH;
    double* work = nullptr;
    if (int st = scratch_on_current_device(&work)) return st;
    return vxm::launch_stamp(tp.dev, batch, params, z, t, cf, cq, gmin, hist, data, res, work,
                             S(stream));
  } catch (...) {
    return VXM_ERR_INVALID;
  }
}
