// fmm_tiles.hip -- the M2L offset groups and target tiles built on the device.
//
// The same partition as fmm::build_m2l_groups and the same tiles, slices and
// work items as fmm::build_m2l_tiles (fmm_host.cpp), for free-space trees
// whose cells all have the radius of their level (trees built by
// fmm::build_tree or fmmdev::device_tree):
//
//  * every M2L pair's offset, in units of its finer cell's size, reduced by
//    common factors of two (the host's lattice key), with (by level) the two
//    levels, packed into one 64-bit key; a hash table on the device gives
//    each pair its key's slot;
//  * on the host, over the distinct keys only (thousands): group numbers in
//    key order, offsets, reflection classes and their representatives --
//    the host's numbering is first-appearance order, which only permutes the
//    groups and classes;
//  * tiles: the targets sorted by (level, octant in the parent, cell), TT per
//    tile -- the host's order, so the same tiles; each tile's pairs sorted by
//    (class, target) and every (tile, class) run of m pairs dealt into S =
//    max(ceil(m / 16), most pairs of one target) slices of distinct targets,
//    pair i of the run to slice i mod S, as the host does.
//
// Keys that do not fit (an offset component of 2^13 or more after the
// reduction, an offset off the lattice) make the whole plan fall back to
// the host; so do periodic images.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <utility>
#include <vector>

#include "fmm_device.hpp"
#include "fmm_host.hpp"

namespace fmmdev {

namespace {

constexpr unsigned long long kEmptyKey = ~0ull;
constexpr int kKBits = 14, kKOff = 1 << (kKBits - 1);  // offset components in [-2^13, 2^13)
constexpr int kHashBits = 20;

struct DBufT {  // device scratch from the plan builders' per-stream cache
  hipStream_t s;
  void *p = nullptr;
  size_t bytes = 0;  // capacity
  explicit DBufT(hipStream_t st) : s(st) {}
  DBufT(const DBufT &) = delete;
  ~DBufT() {
    if (p) plan_scratch_put(p, bytes, s);
  }
  hipError_t need(size_t b) {
    if (b <= bytes) return hipSuccess;
    if (p) plan_scratch_put(p, bytes, s);
    p = nullptr;
    bytes = 0;
    return plan_scratch_get(&p, &bytes, b, s);
  }
  template <class T> T *as() const { return static_cast<T *>(p); }
};

#define TCHK(expr)                   \
  do {                               \
    hipError_t e_ = (expr);          \
    if (e_ != hipSuccess) return e_; \
  } while (0)

__device__ __forceinline__ unsigned long long mix64(unsigned long long k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}

// the target cell of every pair
__global__ void pair_target_kernel(int ncell, const int *__restrict__ ptr, int *__restrict__ ptgt) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  for (int p = ptr[c]; p < ptr[c + 1]; p++) ptgt[p] = c;
}

// every pair's key (host: build_m2l_groups), inserted into the hash table;
// slot[p] its slot, *bad set when a key does not fit
__global__ void pair_key_kernel(int64_t np, const int *__restrict__ ptgt, const int *__restrict__ psrc,
                                const double *__restrict__ X, const int *__restrict__ level, double R0, int by_level,
                                unsigned long long *__restrict__ table, int *__restrict__ slot,
                                int *__restrict__ bad) {
#pragma clang fp contract(off)
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const int c = ptgt[p], sc = psrc[p];
  const int lt = level[c], ls = level[sc];
  int e = lt > ls ? lt : ls;
  const double scale = ldexp(1.0, e) / R0;
  long long k[3];
  bool ok = true;
  for (int d = 0; d < 3; d++) {
    const double v = (X[3 * (size_t)c + d] - X[3 * (size_t)sc + d]) * scale;
    const double rv = rint(v);
    if (!(fabs(v - rv) < 1e-6) || fabs(rv) >= (double)(1 << 26)) ok = false;
    k[d] = (long long)rv;
  }
  if (ok) {
    while (e > 0 && ((k[0] | k[1] | k[2]) & 1) == 0) {
      k[0] /= 2;
      k[1] /= 2;
      k[2] /= 2;
      e--;
    }
    for (int d = 0; d < 3; d++)
      if (k[d] < -kKOff || k[d] >= kKOff) ok = false;
  }
  if (!ok) {
    atomicOr(bad, 1);
    slot[p] = -1;
    return;
  }
  const unsigned long long lv = by_level ? ((unsigned long long)lt << 5 | (unsigned long long)ls) : 0ull;
  const unsigned long long key = (((unsigned long long)e << 10 | lv) << (3 * kKBits)) |
                                 ((unsigned long long)(k[0] + kKOff) << (2 * kKBits)) |
                                 ((unsigned long long)(k[1] + kKOff) << kKBits) | (unsigned long long)(k[2] + kKOff);
  const unsigned long long mask = (1ull << kHashBits) - 1;
  unsigned long long h = mix64(key) & mask;
  for (int probe = 0; probe <= (int)mask; probe++) {
    // a slot changes once, empty -> key: a plain read that shows a key is
    // final (most pairs find theirs so, without an atomic on a hot slot);
    // one that shows empty may be stale, and the CAS decides
    unsigned long long prev = table[h];
    if (prev == kEmptyKey) prev = atomicCAS(&table[h], kEmptyKey, key);
    if (prev == kEmptyKey || prev == key) {
      slot[p] = (int)h;
      return;
    }
    h = (h + 1) & mask;
  }
  atomicOr(bad, 2);  // table full
  slot[p] = -1;
}

// the filled slots of the hash table (any order: the host sorts them)
__global__ void table_compact_kernel(int tsize, const unsigned long long *__restrict__ table,
                                     unsigned long long *__restrict__ okey, int *__restrict__ oslot,
                                     int *__restrict__ count) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= tsize) return;
  const unsigned long long k = table[h];
  if (k == kEmptyKey) return;
  const int at = atomicAdd(count, 1);
  okey[at] = k;
  oslot[at] = h;
}

// slot2grp[gslot[g]] = g: the group of every filled slot
__global__ void slot_group_kernel(int ng, const int *__restrict__ gslot, int *__restrict__ slot2grp) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g < ng) slot2grp[gslot[g]] = g;
}

// the octant of every target cell in its parent; sort keys (level, octant, cell)
__global__ void target_key_kernel(int ncell, const int *__restrict__ ptr, const int *__restrict__ level,
                                  const int *__restrict__ iparent, const double *__restrict__ X,
                                  unsigned long long *__restrict__ key, int *__restrict__ cell, int *__restrict__ cnt) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell || ptr[c + 1] == ptr[c]) return;
  int o = 0;
  if (c > 0) {
    const int p = iparent[c];
    for (int d = 0; d < 3; d++)
      if (X[3 * (size_t)c + d] > X[3 * (size_t)p + d]) o |= 1 << d;
  }
  const int i = atomicAdd(cnt, 1);
  key[i] = ((unsigned long long)level[c] << 40) | ((unsigned long long)o << 32) | (unsigned)c;
  cell[i] = c;
}

__global__ void tile_of_kernel(int nt, const int *__restrict__ cells, int TT, int *__restrict__ tile_of,
                               int *__restrict__ local) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nt) return;
  tile_of[cells[i]] = i / TT;
  local[cells[i]] = i % TT;
}

// pair sort keys (tile, class, target in the tile); values the pair index
__global__ void pair_tile_key_kernel(int64_t np, const int *__restrict__ ptgt, const int *__restrict__ slot,
                                     const int *__restrict__ slot2grp, const int *__restrict__ gcanon,
                                     const int *__restrict__ tile_of, const int *__restrict__ local,
                                     unsigned long long *__restrict__ key, int *__restrict__ val) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  const int c = ptgt[p], g = slot2grp[slot[p]];
  key[p] = ((unsigned long long)tile_of[c] << 40) | ((unsigned long long)gcanon[g] << 8) | (unsigned)local[c];
  val[p] = (int)p;
}

// sorted position i: flags of run starts (segment = (tile, class), run = target)
__global__ void seg_flags_kernel(int64_t np, const unsigned long long *__restrict__ key, int *__restrict__ segf,
                                 int *__restrict__ runs) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const bool segb = i == 0 || (key[i] >> 8) != (key[i - 1] >> 8);
  const bool runb = i == 0 || key[i] != key[i - 1];
  segf[i] = segb ? 1 : 0;
  runs[i] = runb ? (int)i : 0;  // max-scanned into the run start
}

// per segment: start, length and the longest run of one target
__global__ void seg_stats_kernel(int64_t np, const unsigned long long *__restrict__ key,
                                 const int *__restrict__ segid, const int *__restrict__ runstart,
                                 int *__restrict__ sstart, int *__restrict__ slen, int *__restrict__ smost) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const int sg = segid[i] - 1;
  if (i == 0 || (key[i] >> 8) != (key[i - 1] >> 8)) sstart[sg] = (int)i;
  if (i + 1 == np || (key[i + 1] >> 8) != (key[i] >> 8)) slen[sg] = (int)(i + 1);  // end, made a length below
  // the longest run of one target; runs of one pair cannot raise the slice
  // count max(ceil(m / SW), longest run) above ceil(m / SW) >= 1, so only
  // longer runs (a target's several reflections of one class) take an atomic
  if (i + 1 == np || key[i + 1] != key[i]) {
    const int len = (int)(i - runstart[i] + 1);
    if (len > 1) atomicMax(&smost[sg], len);
  }
}

__global__ void seg_slices_kernel(int nseg, const int *__restrict__ sstart, int *__restrict__ slen,
                                  const int *__restrict__ smost, int SW, int *__restrict__ nsl) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nseg) return;
  const int m = slen[g] - sstart[g];
  slen[g] = m;
  const int a = (m + SW - 1) / SW;
  nsl[g] = a > smost[g] ? a : smost[g];
}

// deal every sorted pair into its slice: the tile's source and (target | reflection << 8)
__global__ void deal_kernel(int64_t np, const unsigned long long *__restrict__ key, const int *__restrict__ val,
                            const int *__restrict__ segid, const int *__restrict__ sstart, const int *__restrict__ slen,
                            const int *__restrict__ nsl, const int *__restrict__ psrc, const int *__restrict__ slot,
                            const int *__restrict__ slot2grp, const unsigned char *__restrict__ grefl,
                            int *__restrict__ tsrc, unsigned short *__restrict__ tlt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const int sg = segid[i] - 1, s0 = sstart[sg], m = slen[sg], S = nsl[sg];
  const int r = (int)(i - s0), j = r % S, q = r / S;
  const int out = s0 + j * (m / S) + min(j, m % S) + q;
  const int p = val[i];
  tsrc[out] = psrc[p];
  tlt[out] = (unsigned short)((key[i] & 255) | ((unsigned)grefl[slot2grp[slot[p]]] << 8));
}

// the slice descriptors of every segment, and each tile's first slice
__global__ void slice_desc_kernel(int nseg, const unsigned long long *__restrict__ key, const int *__restrict__ sstart,
                                  const int *__restrict__ slen, const int *__restrict__ nsl,
                                  const int *__restrict__ sbase, int *__restrict__ slice, int *__restrict__ tptr) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nseg) return;
  const int s0 = sstart[g], m = slen[g], S = nsl[g];
  const unsigned long long k = key[s0];
  const int rep = (int)((k >> 8) & 0xffffffffull), tile = (int)(k >> 40);
  if (g == 0 || (key[sstart[g - 1]] >> 40) != (k >> 40)) tptr[tile] = sbase[g];
  for (int j = 0; j < S; j++) {
    int *o = slice + 4 * (size_t)(sbase[g] + j);
    o[0] = rep;
    o[1] = s0 + j * (m / S) + min(j, m % S);
    o[2] = m / S + (j < m % S ? 1 : 0);
    o[3] = 0;
  }
}

}  // namespace

hipError_t device_groups_tiles(const fmm::Tree &t, const fmm::Lists &L, bool by_level, int TT, int SW, int seg,
                               fmm::M2LGroups &G, fmm::M2LTiles &T, bool &done, hipStream_t s,
                               PlanArrays *keep, const int *dsrc_in) {
  done = false;
  if (keep) {
    keep->tiles = false;
    keep->t_slice.reset();
    keep->t_src.reset();
    keep->t_lt.reset();
    keep->nslice = 0;
  }
  const int nc = t.ncell;
  const int64_t np = L.nm2l;
  if (nc == 0 || np == 0 || np > 0x7fffffff || TT > 255) return hipSuccess;
  for (int c = 0; c < nc; c++)  // regular tree: every radius R0 / 2^level
    if (t.R[c] != t.R[0] / (1 << t.level[c])) return hipSuccess;
  const double R0 = t.R[0];
  DBufT dX(s), dlev(s), dpar(s), dptr(s), dsrc(s), dtgt(s), dtab(s), dslot(s), dbad(s);
  TCHK(dX.need(sizeof(double) * 3 * (size_t)nc));
  TCHK(dlev.need(4 * (size_t)nc));
  TCHK(dpar.need(4 * (size_t)nc));
  TCHK(dptr.need(4 * ((size_t)nc + 1)));
  // the pair sources: given on the device, or the lists' device column when
  // the plan keeps one, else uploaded from L
  const bool dev_src = dsrc_in || (keep && keep->lists && keep->m2l_src.bytes == 4 * (size_t)np);
  if (!dev_src) TCHK(dsrc.need(4 * (size_t)np));
  const int *psrc = dsrc_in ? dsrc_in : dev_src ? keep->m2l_src.as<int>() : dsrc.as<int>();
  TCHK(dtgt.need(4 * (size_t)np));
  TCHK(hipMemcpyAsync(dX.p, t.X.data(), sizeof(double) * 3 * (size_t)nc, hipMemcpyHostToDevice, s));
  TCHK(hipMemcpyAsync(dlev.p, t.level.data(), 4 * (size_t)nc, hipMemcpyHostToDevice, s));
  TCHK(hipMemcpyAsync(dpar.p, t.iparent.data(), 4 * (size_t)nc, hipMemcpyHostToDevice, s));
  TCHK(hipMemcpyAsync(dptr.p, L.m2l_ptr.data(), 4 * ((size_t)nc + 1), hipMemcpyHostToDevice, s));
  if (!dev_src) TCHK(hipMemcpyAsync(dsrc.p, L.m2l_src.data(), 4 * (size_t)np, hipMemcpyHostToDevice, s));
  const unsigned nbc = (unsigned)((nc + 255) / 256), nbp = (unsigned)((np + 255) / 256);
  pair_target_kernel<<<nbc, 256, 0, s>>>(nc, dptr.as<int>(), dtgt.as<int>());
  TCHK(hipGetLastError());
  // keys and the hash table
  const size_t tsize = (size_t)1 << kHashBits;
  TCHK(dtab.need(8 * tsize));
  TCHK(dslot.need(4 * (size_t)np));
  TCHK(dbad.need(4));
  TCHK(hipMemsetAsync(dtab.p, 0xff, 8 * tsize, s));
  TCHK(hipMemsetAsync(dbad.p, 0, 4, s));
  pair_key_kernel<<<nbp, 256, 0, s>>>(np, dtgt.as<int>(), psrc, dX.as<double>(), dlev.as<int>(), R0,
                                       by_level ? 1 : 0, dtab.as<unsigned long long>(), dslot.as<int>(),
                                       dbad.as<int>());
  TCHK(hipGetLastError());
  // the filled slots come back (key, slot), not the whole table
  DBufT dck(s), dcs(s), dcn(s);
  TCHK(dck.need(8 * tsize));
  TCHK(dcs.need(4 * tsize));
  TCHK(dcn.need(4));
  TCHK(hipMemsetAsync(dcn.p, 0, 4, s));
  table_compact_kernel<<<(unsigned)((tsize + 255) / 256), 256, 0, s>>>((int)tsize, dtab.as<unsigned long long>(),
                                                                        dck.as<unsigned long long>(), dcs.as<int>(),
                                                                        dcn.as<int>());
  TCHK(hipGetLastError());
  int bad = 0, nfill = 0;
  TCHK(hipMemcpyAsync(&bad, dbad.p, 4, hipMemcpyDeviceToHost, s));
  TCHK(hipMemcpyAsync(&nfill, dcn.p, 4, hipMemcpyDeviceToHost, s));
  TCHK(hipStreamSynchronize(s));
  if (bad) return hipSuccess;  // the host groups this plan
  std::vector<unsigned long long> fk(nfill);
  std::vector<int> fs(nfill);
  if (nfill) {
    TCHK(hipMemcpyAsync(fk.data(), dck.p, 8 * (size_t)nfill, hipMemcpyDeviceToHost, s));
    TCHK(hipMemcpyAsync(fs.data(), dcs.p, 4 * (size_t)nfill, hipMemcpyDeviceToHost, s));
    TCHK(hipStreamSynchronize(s));
  }
  // groups in key order, classes in representative-key order (host)
  std::vector<std::pair<unsigned long long, int>> keys(nfill);
  for (int f = 0; f < nfill; f++) keys[f] = {fk[f], fs[f]};
  std::sort(keys.begin(), keys.end());
  const int ng = (int)keys.size();
  std::vector<int> gslot(ng), gcanon(ng);
  std::vector<unsigned char> grefl(ng);
  G = fmm::M2LGroups();
  G.ngroup = ng;
  G.dx.resize(3 * (size_t)ng);
  G.rest_ptr.assign((size_t)nc + 1, 0);
  G.grp_ptr.assign(1, 0);
  std::map<unsigned long long, int> cls;  // representative key -> class
  std::vector<std::pair<unsigned long long, int>> gcls(ng);
  const unsigned long long kmask = (1ull << kKBits) - 1;
  for (int g = 0; g < ng; g++) {
    const unsigned long long key = keys[g].first;
    gslot[g] = keys[g].second;
    const int e = (int)(key >> (3 * kKBits + 10));
    const unsigned long long lv = (key >> (3 * kKBits)) & 1023;
    long long k[3];
    for (int d = 0; d < 3; d++) k[d] = (long long)((key >> (kKBits * (2 - d))) & kmask) - kKOff;
    const double u = R0 / std::ldexp(1.0, e);
    unsigned char r = 0;
    unsigned long long ck = (unsigned long long)e << 10 | lv;
    for (int d = 0; d < 3; d++) {
      G.dx[3 * (size_t)g + d] = (double)k[d] * u;
      if (k[d] < 0) r |= (unsigned char)(1 << d);
      ck = ck << kKBits | (unsigned long long)(std::llabs(k[d]) + kKOff);
    }
    grefl[g] = r;
    gcls[g] = {ck, g};
  }
  {
    std::vector<std::pair<unsigned long long, int>> sorted(gcls);
    std::sort(sorted.begin(), sorted.end());
    for (const auto &x : sorted)
      if (!cls.count(x.first)) {
        const int id = (int)cls.size();
        cls[x.first] = id;
        const int g = x.second;
        for (int d = 0; d < 3; d++) G.cdx.push_back(std::fabs(G.dx[3 * (size_t)g + d]));
        if (by_level) {
          const unsigned long long lv = (x.first >> (3 * kKBits)) & 1023;
          G.cscale.insert(G.cscale.end(), {R0 / (1 << (int)(lv & 31)), R0 / (1 << (int)(lv >> 5))});
        }
      }
    for (int g = 0; g < ng; g++) gcanon[g] = cls[gcls[g].first];
  }
  G.ncanon = (int)cls.size();
  G.canon = gcanon;
  G.refl.assign(grefl.begin(), grefl.end());
  if (by_level) {
    G.scale.resize(2 * (size_t)ng);
    for (int g = 0; g < ng; g++) {
      const unsigned long long lv = (keys[g].first >> (3 * kKBits)) & 1023;
      G.scale[2 * (size_t)g] = R0 / (1 << (int)(lv & 31));      // R_s
      G.scale[2 * (size_t)g + 1] = R0 / (1 << (int)(lv >> 5));  // R_t
    }
  }
  // tiles: targets in (level, octant, cell) order
  DBufT dtk(s), dtk2(s), dtc(s), dtc2(s), dcnt(s), dtof(s), dloc(s), tmp(s);
  TCHK(dtk.need(8 * (size_t)nc));
  TCHK(dtk2.need(8 * (size_t)nc));
  TCHK(dtc.need(4 * (size_t)nc));
  TCHK(dtc2.need(4 * (size_t)nc));
  TCHK(dcnt.need(4));
  TCHK(dtof.need(4 * (size_t)nc));
  TCHK(dloc.need(4 * (size_t)nc));
  TCHK(hipMemsetAsync(dcnt.p, 0, 4, s));
  target_key_kernel<<<nbc, 256, 0, s>>>(nc, dptr.as<int>(), dlev.as<int>(), dpar.as<int>(), dX.as<double>(),
                                        dtk.as<unsigned long long>(), dtc.as<int>(), dcnt.as<int>());
  TCHK(hipGetLastError());
  int ntg = 0;
  TCHK(hipMemcpyAsync(&ntg, dcnt.p, 4, hipMemcpyDeviceToHost, s));
  TCHK(hipStreamSynchronize(s));
  size_t tb = 0;
  TCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, dtk.as<unsigned long long>(), dtk2.as<unsigned long long>(),
                                          dtc.as<int>(), dtc2.as<int>(), ntg, 0, 48, s));
  TCHK(tmp.need(tb));
  TCHK(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, dtk.as<unsigned long long>(), dtk2.as<unsigned long long>(),
                                          dtc.as<int>(), dtc2.as<int>(), ntg, 0, 48, s));
  tile_of_kernel<<<(unsigned)((ntg + 255) / 256), 256, 0, s>>>(ntg, dtc2.as<int>(), TT, dtof.as<int>(),
                                                                dloc.as<int>());
  TCHK(hipGetLastError());
  T = fmm::M2LTiles();
  T.TT = TT;
  T.ntile = (ntg + TT - 1) / TT;
  T.tgt.assign((size_t)T.ntile * TT, -1);
  TCHK(hipMemcpyAsync(T.tgt.data(), dtc2.p, 4 * (size_t)ntg, hipMemcpyDeviceToHost, s));
  // pairs sorted by (tile, class, target)
  DBufT ds2g(s), dgc(s), dgr(s), dpk(s), dpk2(s), dpv(s), dpv2(s);
  TCHK(ds2g.need(4 * tsize));
  TCHK(dgc.need(4 * (size_t)ng));
  TCHK(dgr.need((size_t)ng));
  {  // every filled slot's group, scattered on the device
    DBufT dgs(s);
    TCHK(dgs.need(4 * (size_t)ng));
    TCHK(hipMemcpyAsync(dgs.p, gslot.data(), 4 * (size_t)ng, hipMemcpyHostToDevice, s));
    slot_group_kernel<<<(unsigned)((ng + 255) / 256), 256, 0, s>>>(ng, dgs.as<int>(), ds2g.as<int>());
    TCHK(hipGetLastError());
  }
  TCHK(hipMemcpyAsync(dgc.p, gcanon.data(), 4 * (size_t)ng, hipMemcpyHostToDevice, s));
  TCHK(hipMemcpyAsync(dgr.p, grefl.data(), (size_t)ng, hipMemcpyHostToDevice, s));
  TCHK(dpk.need(8 * (size_t)np));
  TCHK(dpk2.need(8 * (size_t)np));
  TCHK(dpv.need(4 * (size_t)np));
  TCHK(dpv2.need(4 * (size_t)np));
  pair_tile_key_kernel<<<nbp, 256, 0, s>>>(np, dtgt.as<int>(), dslot.as<int>(), ds2g.as<int>(), dgc.as<int>(),
                                           dtof.as<int>(), dloc.as<int>(), dpk.as<unsigned long long>(),
                                           dpv.as<int>());
  TCHK(hipGetLastError());
  int tbits = 1;
  while ((1 << tbits) < T.ntile) tbits++;
  size_t tb2 = 0;
  TCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb2, dpk.as<unsigned long long>(), dpk2.as<unsigned long long>(),
                                          dpv.as<int>(), dpv2.as<int>(), (int)np, 0, 40 + tbits, s));
  TCHK(tmp.need(tb2));
  TCHK(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb2, dpk.as<unsigned long long>(), dpk2.as<unsigned long long>(),
                                          dpv.as<int>(), dpv2.as<int>(), (int)np, 0, 40 + tbits, s));
  // segments (tile, class) and runs (target)
  DBufT dsegf(s), dsegid(s), druns(s), drs(s);
  TCHK(dsegf.need(4 * (size_t)np));
  TCHK(dsegid.need(4 * (size_t)np));
  TCHK(druns.need(4 * (size_t)np));
  TCHK(drs.need(4 * (size_t)np));
  seg_flags_kernel<<<nbp, 256, 0, s>>>(np, dpk2.as<unsigned long long>(), dsegf.as<int>(), druns.as<int>());
  TCHK(hipGetLastError());
  size_t ts1 = 0, ts2 = 0;
  TCHK(hipcub::DeviceScan::InclusiveSum(nullptr, ts1, dsegf.as<int>(), dsegid.as<int>(), (int)np, s));
  TCHK(hipcub::DeviceScan::InclusiveScan(nullptr, ts2, druns.as<int>(), drs.as<int>(), hipcub::Max(), (int)np, s));
  TCHK(tmp.need(std::max(ts1, ts2)));
  TCHK(hipcub::DeviceScan::InclusiveSum(tmp.p, ts1, dsegf.as<int>(), dsegid.as<int>(), (int)np, s));
  TCHK(hipcub::DeviceScan::InclusiveScan(tmp.p, ts2, druns.as<int>(), drs.as<int>(), hipcub::Max(), (int)np, s));
  int nseg = 0;
  TCHK(hipMemcpyAsync(&nseg, dsegid.as<int>() + (np - 1), 4, hipMemcpyDeviceToHost, s));
  TCHK(hipStreamSynchronize(s));
  DBufT dss(s), dsl(s), dsm(s), dns(s), dsb(s);
  for (DBufT *b : {&dss, &dsl, &dsm, &dns, &dsb}) TCHK(b->need(4 * ((size_t)nseg + 1)));
  TCHK(hipMemsetAsync(dsm.p, 0, 4 * (size_t)nseg, s));
  seg_stats_kernel<<<nbp, 256, 0, s>>>(np, dpk2.as<unsigned long long>(), dsegid.as<int>(), drs.as<int>(),
                                       dss.as<int>(), dsl.as<int>(), dsm.as<int>());
  TCHK(hipGetLastError());
  const unsigned nbs = (unsigned)((nseg + 255) / 256);
  seg_slices_kernel<<<nbs, 256, 0, s>>>(nseg, dss.as<int>(), dsl.as<int>(), dsm.as<int>(), SW, dns.as<int>());
  TCHK(hipGetLastError());
  size_t ts3 = 0;
  TCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, ts3, dns.as<int>(), dsb.as<int>(), nseg + 1, s));
  TCHK(tmp.need(ts3));
  TCHK(hipMemsetAsync(dns.as<int>() + nseg, 0, 4, s));
  TCHK(hipcub::DeviceScan::ExclusiveSum(tmp.p, ts3, dns.as<int>(), dsb.as<int>(), nseg + 1, s));
  int nslice = 0;
  TCHK(hipMemcpyAsync(&nslice, dsb.as<int>() + nseg, 4, hipMemcpyDeviceToHost, s));
  TCHK(hipStreamSynchronize(s));
  // the tiles' pair sources, pair targets and slices: on the device in keep,
  // else staged here and copied to T
  DBufT dts(s), dtl(s), dsd(s), dtp(s);
  int *osrc, *osd;
  unsigned short *olt;
  if (keep) {
    int dev = 0;
    TCHK(hipGetDevice(&dev));
    TCHK(keep->t_src.alloc(4 * (size_t)np, dev));
    TCHK(keep->t_lt.alloc(2 * (size_t)np, dev));
    TCHK(keep->t_slice.alloc(16 * (size_t)nslice, dev));
    osrc = keep->t_src.as<int>();
    olt = keep->t_lt.as<unsigned short>();
    osd = keep->t_slice.as<int>();
  } else {
    TCHK(dts.need(4 * (size_t)np));
    TCHK(dtl.need(2 * (size_t)np));
    TCHK(dsd.need(16 * (size_t)nslice));
    osrc = dts.as<int>();
    olt = dtl.as<unsigned short>();
    osd = dsd.as<int>();
  }
  TCHK(dtp.need(4 * ((size_t)T.ntile + 1)));
  deal_kernel<<<nbp, 256, 0, s>>>(np, dpk2.as<unsigned long long>(), dpv2.as<int>(), dsegid.as<int>(), dss.as<int>(),
                                  dsl.as<int>(), dns.as<int>(), psrc, dslot.as<int>(), ds2g.as<int>(),
                                  dgr.as<unsigned char>(), osrc, olt);
  TCHK(hipGetLastError());
  slice_desc_kernel<<<nbs, 256, 0, s>>>(nseg, dpk2.as<unsigned long long>(), dss.as<int>(), dsl.as<int>(),
                                        dns.as<int>(), dsb.as<int>(), osd, dtp.as<int>());
  TCHK(hipGetLastError());
  T.ptr.resize((size_t)T.ntile + 1);
  TCHK(hipMemcpyAsync(T.ptr.data(), dtp.p, 4 * (size_t)T.ntile, hipMemcpyDeviceToHost, s));
  if (keep) {
    T.slice.clear();
    T.src.clear();
    T.lt.clear();
  } else {
    T.slice.resize(4 * (size_t)nslice);
    T.src.resize(np);
    T.lt.resize(np);
    TCHK(hipMemcpyAsync(T.slice.data(), dsd.p, 16 * (size_t)nslice, hipMemcpyDeviceToHost, s));
    TCHK(hipMemcpyAsync(T.src.data(), dts.p, 4 * (size_t)np, hipMemcpyDeviceToHost, s));
    TCHK(hipMemcpyAsync(T.lt.data(), dtl.p, 2 * (size_t)np, hipMemcpyDeviceToHost, s));
  }
  TCHK(hipStreamSynchronize(s));
  T.ptr[T.ntile] = nslice;
  if (keep) {
    TCHK(hipGetDevice(&keep->device));
    keep->nslice = nslice;
    keep->tiles = true;
  }
  fmm::split_m2l_tiles(T, seg);
  done = true;
  return hipSuccess;
}

}  // namespace fmmdev
