// lu_plan.cpp -- turn the symbolic analysis into device-resident maps for
// the multifrontal factorization and solves.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "lu.hpp"

namespace tailor {

namespace {

template <class T>
hipError_t upload(T **dst, const std::vector<T> &src, hipStream_t s) {
  hipError_t e = hipMalloc(reinterpret_cast<void **>(dst), std::max<size_t>(1, src.size()) * sizeof(T));
  if (e != hipSuccess) return e;
  if (src.empty()) return hipSuccess;
  return hipMemcpyAsync(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, s);
}

template <class T>
void release(T *&p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

// The look-ahead stream: non-blocking, on every compute unit but the last
// kSideReservedCus; a plain non-blocking stream where the device has too few
// CUs or refuses the mask.
hipError_t side_stream(hipStream_t *s) {
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
      ncu > 4 * kSideReservedCus) {
    std::vector<uint32_t> mask((ncu + 31) / 32, 0);
    for (int c = 0; c < ncu - kSideReservedCus; ++c) mask[c / 32] |= 1u << (c % 32);
    if (hipExtStreamCreateWithCUMask(s, static_cast<uint32_t>(mask.size()), mask.data()) ==
        hipSuccess)
      return hipSuccess;
    (void)hipGetLastError();
  }
  return hipStreamCreateWithFlags(s, hipStreamNonBlocking);
}

}  // namespace

hipError_t lu_plan_create(const Symbolic &S, const std::vector<int> &colidx, int per, LuPlan &P,
                          hipStream_t s) {
  lu_plan_destroy(P);
  P.nsn = S.nsn;
  P.nb = kLuNb;
  P.n = 5LL * S.nodes;
  std::vector<FrontDev> fr(S.nsn);
  long long off = 0, woff = 0;
  int poff = 0;
  for (int f = 0; f < S.nsn; ++f) {
    FrontDev &d = fr[f];
    d.m = (S.m[f] + kFrontAlign - 1) / kFrontAlign * kFrontAlign;
    d.p = S.p[f];
    d.c = S.c[f];
    d.pe = (d.p + kFrontAlign - 1) / kFrontAlign * kFrontAlign;
    off = (off + kFrontAlign - 1) / kFrontAlign * kFrontAlign;
    d.off = off;
    d.woff = woff;
    d.piv_off = poff;
    d.dof0 = 5 * S.sn_start[f];
    d.dinv_off = 0;
    d.inst = 0;
    off += 2LL * d.m * d.m + 4LL * d.pe * d.p;
    woff += d.m;
    poff += d.p;
  }
  P.arena_doubles = off;
  P.w_len = woff;

  // level lists
  std::vector<int> level_list;
  P.levels.assign(S.levels.size(), {});
  for (size_t l = 0; l < S.levels.size(); ++l) {
    LevelPlan &L = P.levels[l];
    L.list_off = static_cast<int>(level_list.size());
    L.count = static_cast<int>(S.levels[l].size());
    for (int f : S.levels[l]) {
      level_list.push_back(f);
      L.pmax = std::max(L.pmax, S.p[f]);
      L.cmax = std::max(L.cmax, S.c[f]);
      L.mmax = std::max(L.mmax, fr[f].m);
    }
    // inverse workspace: the level's fronts take consecutive panel slots
    long long slots = 0;
    for (int f : S.levels[l]) {
      fr[f].dinv_off = static_cast<int>(slots);
      slots += (S.p[f] + kLuNb - 1) / kLuNb;
    }
    P.dinv_slots = std::max(P.dinv_slots, slots);
    P.count_max = std::max(P.count_max, L.count);
  }
  std::vector<int> parent = S.parent, rel, child_ptr(S.nsn + 1, 0), child_list, upd_pos;
  std::vector<long long> rel_off(S.nsn, 0), upd_off(S.nsn, 0);
  // front build: for each front and each of its (at most two) children, the
  // child's update node at each node of the front (the inverse of rel)
  P.h_build.assign(static_cast<size_t>(kBuildInts) * S.nsn, -1);
  P.h_inv.clear();
  for (int f = 0; f < S.nsn; ++f) {
    if (S.children[f].size() > 2) return hipErrorInvalidValue;
    const int nodes = S.m[f] / 5;
    for (size_t k = 0; k < S.children[f].size(); ++k) {
      const int ch = S.children[f][k];
      P.h_build[kBuildInts * f + k] = ch;
      P.h_build[kBuildInts * f + 2 + k] = static_cast<int>(P.h_inv.size());
      const size_t base = P.h_inv.size();
      P.h_inv.resize(base + nodes, -1);
      for (size_t n = 0; n < S.rel[ch].size(); ++n) {
        const int q = S.rel[ch][n];
        if (q < 0 || q >= nodes) return hipErrorInvalidValue;
        P.h_inv[base + q] = static_cast<int>(n);
      }
    }
  }
  for (int f = 0; f < S.nsn; ++f) {
    rel_off[f] = static_cast<long long>(rel.size());
    rel.insert(rel.end(), S.rel[f].begin(), S.rel[f].end());
    child_ptr[f + 1] = child_ptr[f] + static_cast<int>(S.children[f].size());
    child_list.insert(child_list.end(), S.children[f].begin(), S.children[f].end());
    upd_off[f] = static_cast<long long>(upd_pos.size());
    for (int v : S.strct[f]) upd_pos.push_back(S.pos[v]);
  }
  // assembly map: every stored block of A goes to the front of whichever of
  // its row / column node is eliminated first
  std::vector<int> sn_of_pos(S.nodes);
  for (int f = 0; f < S.nsn; ++f)
    for (int q = S.sn_start[f]; q < S.sn_start[f + 1]; ++q) sn_of_pos[q] = f;
  const long long nnzb = static_cast<long long>(colidx.size());
  std::vector<int> amap(3 * nnzb);
  for (long long k = 0; k < nnzb; ++k) {
    const int u = static_cast<int>(k / per), v = colidx[k];
    const int f = sn_of_pos[std::min(S.pos[u], S.pos[v])];
    const int lu = S.local_index(f, u), lv = S.local_index(f, v);
    if (lu < 0 || lv < 0) return hipErrorInvalidValue;
    amap[3 * k] = f;
    amap[3 * k + 1] = lu;
    amap[3 * k + 2] = lv;
  }
  // the blocks grouped by the level of their front (scattered level by
  // level, after the level's fronts are built)
  std::vector<int> ablk;
  ablk.reserve(nnzb);
  {
    std::vector<int> level_of(S.nsn, 0);
    for (size_t l = 0; l < S.levels.size(); ++l)
      for (int f : S.levels[l]) level_of[f] = static_cast<int>(l);
    std::vector<long long> count(S.levels.size() + 1, 0);
    for (long long k = 0; k < nnzb; ++k) ++count[level_of[amap[3 * k]] + 1];
    for (size_t l = 0; l < S.levels.size(); ++l) count[l + 1] += count[l];
    for (size_t l = 0; l < S.levels.size(); ++l) {
      P.levels[l].ablk_off = count[l];
      P.levels[l].ablk_count = count[l + 1] - count[l];
    }
    ablk.assign(nnzb, 0);
    std::vector<long long> next(count.begin(), count.end() - 1);
    for (long long k = 0; k < nnzb; ++k) ablk[next[level_of[amap[3 * k]]]++] = static_cast<int>(k);
  }
  std::vector<int> perm(P.n);
  for (int q = 0; q < S.nodes; ++q)
    for (int dof = 0; dof < 5; ++dof) perm[5 * q + dof] = 5 * S.order[q] + dof;

  hipError_t e;
  if ((e = upload(&P.d_front, fr, s)) != hipSuccess) return e;
  P.h_front = fr;
  P.h_level_list = level_list;
  if ((e = upload(&P.d_level_list, level_list, s)) != hipSuccess) return e;
  if ((e = upload(&P.d_build, P.h_build, s)) != hipSuccess) return e;
  if ((e = upload(&P.d_inv, P.h_inv, s)) != hipSuccess) return e;
  if ((e = upload(&P.d_ablk, ablk, s)) != hipSuccess) return e;
  if ((e = upload(&P.d_parent, parent, s)) != hipSuccess) return e;
  if ((e = upload(&P.d_rel, rel, s)) != hipSuccess) return e;
  if ((e = upload(&P.d_rel_off, rel_off, s)) != hipSuccess) return e;
  if ((e = upload(&P.d_child_ptr, child_ptr, s)) != hipSuccess) return e;
  if ((e = upload(&P.d_child_list, child_list, s)) != hipSuccess) return e;
  if ((e = upload(&P.d_upd_pos, upd_pos, s)) != hipSuccess) return e;
  if ((e = upload(&P.d_upd_off, upd_off, s)) != hipSuccess) return e;
  if ((e = upload(&P.d_amap, amap, s)) != hipSuccess) return e;
  if ((e = upload(&P.d_perm, perm, s)) != hipSuccess) return e;
  std::vector<int> zeros(std::max(1, poff), 0);
  if ((e = upload(&P.d_ipiv, zeros, s)) != hipSuccess) return e;
  if ((e = upload(&P.d_counters, std::vector<int>(8, 0), s)) != hipSuccess) return e;
  if ((e = upload(&P.d_fparams, std::vector<double>(kFpCount, 0.0), s)) != hipSuccess) return e;
  if ((e = hipMalloc(reinterpret_cast<void **>(&P.arena), P.arena_doubles * sizeof(double))) != hipSuccess)
    return e;
  if ((e = hipMalloc(reinterpret_cast<void **>(&P.W), 2 * P.w_len * sizeof(double))) != hipSuccess)
    return e;
  if ((e = hipMalloc(reinterpret_cast<void **>(&P.dinv), std::max<long long>(1, P.dinv_slots) *
                                                             kLuInvSlot * sizeof(double))) !=
      hipSuccess)
    return e;
  if ((e = hipMalloc(reinterpret_cast<void **>(&P.d_bperm),
                     std::max(1, P.count_max) * static_cast<size_t>(kLuPermStride) * sizeof(int))) !=
      hipSuccess)
    return e;
  if ((e = hipMalloc(reinterpret_cast<void **>(&P.Y), 2 * P.n * sizeof(double))) != hipSuccess)
    return e;
  if ((e = hipMalloc(reinterpret_cast<void **>(&P.W2), 2 * P.w_len * sizeof(double))) != hipSuccess)
    return e;
  if ((e = hipMalloc(reinterpret_cast<void **>(&P.Y2), 2 * P.n * sizeof(double))) != hipSuccess)
    return e;
  if ((e = hipEventCreateWithFlags(&P.ev_head, hipEventDisableTiming)) != hipSuccess) return e;
  if ((e = hipEventCreateWithFlags(&P.ev_panel, hipEventDisableTiming)) != hipSuccess) return e;
  return hipStreamSynchronize(s);
}

hipError_t lu_plan_batch(const std::vector<BatchPart> &parts, bool fp32, LuPlan &U,
                         std::vector<int> &fbase, hipStream_t s) {
  lu_plan_destroy(U);
  if (parts.empty()) return hipErrorInvalidValue;
  const size_t unit = fp32 ? sizeof(float) : sizeof(double);
  const char *base = nullptr;
  size_t nlev = 0, deepest = 0;
  for (size_t b = 0; b < parts.size(); ++b) {
    const LuPlan &P = *parts[b].plan;
    if (!P.arena || !parts[b].sym || parts[b].sym->nsn != P.nsn) return hipErrorInvalidValue;
    const char *a = reinterpret_cast<const char *>(P.arena);
    if (!base || a < base) base = a;
    if (P.levels.size() > nlev) {
      nlev = P.levels.size();
      deepest = b;
    }
  }
  // fronts: offsets from the lowest arena in front-type units, pivot offsets
  // after the previous parts' pivots, the part as instance
  std::vector<FrontDev> fr;
  std::vector<int> parent, rel;
  std::vector<long long> rel_off;
  fbase.assign(parts.size(), 0);
  U.h_build.clear();
  U.h_inv.clear();
  int poff = 0;
  for (size_t b = 0; b < parts.size(); ++b) {
    const LuPlan &P = *parts[b].plan;
    const Symbolic &S = *parts[b].sym;
    const long long bytes = reinterpret_cast<const char *>(P.arena) - base;
    if (bytes % static_cast<long long>(unit) != 0) return hipErrorInvalidValue;
    const long long delta = bytes / static_cast<long long>(unit);
    const int fb = static_cast<int>(fr.size());
    fbase[b] = fb;
    for (int f = 0; f < P.nsn; ++f) {
      FrontDev d = P.h_front[f];
      d.off += delta;
      d.piv_off += poff;
      d.inst = static_cast<int>(b);
      fr.push_back(d);
      parent.push_back(S.parent[f] < 0 ? -1 : S.parent[f] + fb);
      rel_off.push_back(static_cast<long long>(rel.size()));
      rel.insert(rel.end(), S.rel[f].begin(), S.rel[f].end());
    }
    // the part's build records with batch front indices and map offsets
    if (P.h_build.size() != static_cast<size_t>(kBuildInts) * P.nsn) return hipErrorInvalidValue;
    const int ib = static_cast<int>(U.h_inv.size());
    for (int f = 0; f < P.nsn; ++f) {
      const int *rec = &P.h_build[kBuildInts * f];
      for (int k = 0; k < 2; ++k) U.h_build.push_back(rec[k] < 0 ? -1 : rec[k] + fb);
      for (int k = 0; k < 2; ++k) U.h_build.push_back(rec[k] < 0 ? -1 : rec[2 + k] + ib);
    }
    U.h_inv.insert(U.h_inv.end(), P.h_inv.begin(), P.h_inv.end());
    for (int f = 0; f < P.nsn; ++f) poff += P.h_front[f].p;
  }
  // level l of the batch: level l of every part, part by part
  std::vector<int> level_list;
  U.levels.assign(nlev, {});
  for (size_t l = 0; l < nlev; ++l) {
    LevelPlan &L = U.levels[l];
    L.list_off = static_cast<int>(level_list.size());
    for (size_t b = 0; b < parts.size(); ++b) {
      const LuPlan &P = *parts[b].plan;
      if (l >= P.levels.size()) continue;
      const LevelPlan &Lb = P.levels[l];
      for (int q = 0; q < Lb.count; ++q)
        level_list.push_back(P.h_level_list[Lb.list_off + q] + fbase[b]);
    }
    L.count = static_cast<int>(level_list.size()) - L.list_off;
    long long slots = 0;
    for (int q = L.list_off; q < L.list_off + L.count; ++q) {
      FrontDev &d = fr[level_list[q]];
      L.pmax = std::max(L.pmax, d.p);
      L.cmax = std::max(L.cmax, d.c);
      L.mmax = std::max(L.mmax, d.m);
      d.dinv_off = static_cast<int>(slots);
      slots += (d.p + kLuNb - 1) / kLuNb;
    }
    U.dinv_slots = std::max(U.dinv_slots, slots);
    U.count_max = std::max(U.count_max, L.count);
  }
  U.nsn = static_cast<int>(fr.size());
  U.nb = kLuNb;
  U.instances = static_cast<int>(parts.size());
  U.gemm_shape = parts[deepest].plan->gemm_shape;
  U.gemm_order = parts[deepest].plan->gemm_order;
  U.panel_streaming = parts[deepest].plan->panel_streaming;
  U.pair_blocks = parts[deepest].plan->pair_blocks;
  U.h_front = fr;
  U.h_level_list = level_list;
  U.arena = reinterpret_cast<double *>(const_cast<char *>(base));
  U.owns_arena = false;

  hipError_t e;
  if ((e = upload(&U.d_front, fr, s)) != hipSuccess) return e;
  if ((e = upload(&U.d_level_list, level_list, s)) != hipSuccess) return e;
  if ((e = upload(&U.d_build, U.h_build, s)) != hipSuccess) return e;
  if ((e = upload(&U.d_inv, U.h_inv, s)) != hipSuccess) return e;
  if ((e = upload(&U.d_parent, parent, s)) != hipSuccess) return e;
  if ((e = upload(&U.d_rel, rel, s)) != hipSuccess) return e;
  if ((e = upload(&U.d_rel_off, rel_off, s)) != hipSuccess) return e;
  if ((e = upload(&U.d_ipiv, std::vector<int>(std::max(1, poff), 0), s)) != hipSuccess) return e;
  if ((e = upload(&U.d_counters, std::vector<int>(std::max<size_t>(8, parts.size()), 0), s)) !=
      hipSuccess)
    return e;
  if ((e = upload(&U.d_fparams, std::vector<double>(kFpCount * parts.size(), 0.0), s)) !=
      hipSuccess)
    return e;
  if ((e = hipMalloc(reinterpret_cast<void **>(&U.dinv), std::max<long long>(1, U.dinv_slots) *
                                                             kLuInvSlot * sizeof(double))) !=
      hipSuccess)
    return e;
  if ((e = hipMalloc(reinterpret_cast<void **>(&U.d_bperm),
                     std::max(1, U.count_max) * static_cast<size_t>(kLuPermStride) * sizeof(int))) !=
      hipSuccess)
    return e;
  // (batch factorizations replay as HIP graphs, which do not keep a
  // stream's CU mask: a plain stream)
  if ((e = hipStreamCreateWithFlags(&U.side, hipStreamNonBlocking)) != hipSuccess) return e;
  if ((e = hipEventCreateWithFlags(&U.ev_head, hipEventDisableTiming)) != hipSuccess) return e;
  if ((e = hipEventCreateWithFlags(&U.ev_panel, hipEventDisableTiming)) != hipSuccess) return e;
  return hipStreamSynchronize(s);
}

hipError_t lu_plan_side(const LuPlan &P) {
  return P.side ? hipSuccess : side_stream(&P.side);
}

void lu_plan_destroy(LuPlan &P) {
  if (P.side) (void)hipStreamSynchronize(P.side);
  if (P.ev_head) (void)hipEventDestroy(P.ev_head);
  if (P.ev_panel) (void)hipEventDestroy(P.ev_panel);
  if (P.side) (void)hipStreamDestroy(P.side);
  P.side = nullptr;
  P.ev_head = P.ev_panel = nullptr;
  release(P.d_front);
  release(P.d_level_list);
  release(P.d_build);
  release(P.d_inv);
  release(P.d_ablk);
  release(P.d_parent);
  release(P.d_rel);
  release(P.d_rel_off);
  release(P.d_child_ptr);
  release(P.d_child_list);
  release(P.d_upd_pos);
  release(P.d_upd_off);
  release(P.d_amap);
  release(P.d_ipiv);
  release(P.d_counters);
  release(P.d_fparams);
  release(P.d_perm);
  if (P.owns_arena)
    release(P.arena);
  else
    P.arena = nullptr;
  P.owns_arena = true;
  P.instances = 1;
  release(P.W);
  release(P.dinv);
  release(P.d_bperm);
  release(P.Y);
  release(P.W2);
  release(P.Y2);
  P.levels.clear();
  P.dinv_slots = 0;
  P.count_max = 0;
}

}  // namespace tailor
