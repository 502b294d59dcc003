// Node-set builder of the site-class, point-group reduced augmented-space operator.
//
// Follows build_augcls_operator (src/asr_augcls.f90:596-823) with canon_node (:494-526), set_state / get_state (:208-235)
// (src/asr_augspace.f90:48-170 keeps the nodes of the closure in a hashed store; NodeStore below is this builder's own):
//   * breadth-first closure over paths R -> R1 -> R2 with every local-state combination of the visited sites, rank-limited;
//   * every candidate is canonicalised over the group (minimum over the operations of (site, sorted config codes));
//   * candidates are inserted serially in the order (node, k1, k2, f) of the plain serial loop, so node numbering does not depend
//     on the number of threads (the candidates of a chunk of nodes are computed in parallel, as the reference does);
//   * then the per-node tables: stabilisers, flip partners (flipr/flipg), hop links (hopr/hopg), local state, orbit size.
// Indices are 0-based here: sites 0..nsite-1, nodes 0..nn-1, operations 0..nop-1, -1 = no link.  A config code is
// site*nsm + state (state 1..nsc-1); a shift of every site index by one keeps every comparison, so the canonical
// representatives and the node order are those of the 1-based reference.
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <cstdio>
#include <vector>
#include <algorithm>
#include <stdexcept>
#include <string>
#include <omp.h>

#include "asr_host.h"

namespace {

// The node store: every node's key (its site and sorted config codes) in one arena, looked up through an open-addressed
// index.  Node numbers are the insertion order; the index only finds keys, so the hash function does not affect the node
// numbering.  The index keeps the 64-bit hash of each entry beside it (most probes end on a hash mismatch without
// touching the arena), holds at most half as many entries as slots, and doubles when it reaches that.
struct NodeStore {
  std::vector<int> site, start, len, level, orb, data;  // per node; data: the config codes of every node, back to back
  std::vector<uint64_t> khash;                          // per node: hash of its key
  std::vector<int32_t> slot;                            // index: node number or -1
  std::vector<uint64_t> slot_hash;                      // index: hash of the entry in that slot
  size_t mask = 0;
  int n = 0;

  static uint64_t mix(uint64_t h) {  // a 64-bit finaliser (xor-shift and multiply rounds)
    h ^= h >> 31;
    h *= 0x7fb5d329728ea185ull;
    h ^= h >> 27;
    h *= 0x81dadef4bc2dd44dull;
    return h ^ (h >> 33);
  }
  static uint64_t key_hash(int s, const int* cfg, int nc) {
    uint64_t h = mix(0x243f6a8885a308d3ull ^ (uint64_t)(uint32_t)s);
    for (int i = 0; i < nc; ++i) h = mix(h + 0x9e3779b97f4a7c15ull * (uint64_t)((uint32_t)cfg[i] + 1u));
    return h;
  }
  void init(int cap) {
    site.reserve(cap); start.reserve(cap); len.reserve(cap); level.reserve(cap); orb.reserve(cap);
    khash.reserve(cap); data.reserve(8 * (size_t)cap);
    size_t sz = 64;
    while (sz < 2 * (size_t)cap) sz <<= 1;
    slot.assign(sz, -1);
    slot_hash.assign(sz, 0);
    mask = sz - 1;
  }
  bool key_is(int j, int s, const int* cfg, int nc) const {
    if (site[j] != s || len[j] != nc) return false;
    return std::equal(cfg, cfg + nc, data.begin() + start[j]);
  }
  // the slot holding the key, or the empty slot where it would go
  size_t probe(uint64_t h, int s, const int* cfg, int nc) const {
    size_t q = (size_t)h & mask;
    while (slot[q] >= 0 && !(slot_hash[q] == h && key_is(slot[q], s, cfg, nc))) q = (q + 1) & mask;
    return q;
  }
  int find(int s, const int* cfg, int nc) const { return slot[probe(key_hash(s, cfg, nc), s, cfg, nc)]; }
  void grow() {
    const size_t sz = 2 * (mask + 1);
    slot.assign(sz, -1);
    slot_hash.assign(sz, 0);
    mask = sz - 1;
    for (int j = 0; j < n; ++j) {
      size_t q = (size_t)khash[j] & mask;
      while (slot[q] >= 0) q = (q + 1) & mask;
      slot[q] = j;
      slot_hash[q] = khash[j];
    }
  }
  int find_or_add(int s, const int* cfg, int nc, int lev, bool* isnew) {
    const uint64_t h = key_hash(s, cfg, nc);
    const size_t q = probe(h, s, cfg, nc);
    if (slot[q] >= 0) {
      *isnew = false;
      return slot[q];
    }
    const int j = n++;
    site.push_back(s); len.push_back(nc); level.push_back(lev); orb.push_back(1); khash.push_back(h);
    start.push_back((int)data.size());
    data.insert(data.end(), cfg, cfg + nc);
    slot[q] = j;
    slot_hash[q] = h;
    *isnew = true;
    if (2 * (size_t)n > mask) grow();
    return j;
  }
};

inline void sort_small(int* a, int n) {
  for (int i = 1; i < n; ++i) {
    int t = a[i], j = i - 1;
    while (j >= 0 && a[j] > t) { a[j + 1] = a[j]; --j; }
    a[j + 1] = t;
  }
}

// canon_node (src/asr_augcls.f90:494-526): minimum of (perm(site), sorted mapped codes) over the operations;
// gbest = first operation reaching it, osize = nop / (number of operations reaching it).
struct Canon {
  int csite, osize, gbest;
  std::vector<int> ccfg;
};

inline void canon_node(const int* perm, int nsite, int nop, int nsm, int site, const int* cfg, int n, Canon& out,
                       std::vector<int>& c2) {
  out.ccfg.assign(n, 0);
  c2.resize(n);
  out.csite = 0x7fffffff;
  for (int k = 0; k < n; ++k) out.ccfg[k] = 0x7fffffff;
  int ntie = 0;
  out.gbest = 0;
  for (int op = 0; op < nop; ++op) {
    const int* p = perm + (size_t)op * nsite;
    int s2 = p[site];
    for (int k = 0; k < n; ++k) c2[k] = p[cfg[k] / nsm] * nsm + cfg[k] % nsm;
    sort_small(c2.data(), n);
    bool better = false, equal = false;
    if (s2 < out.csite) {
      better = true;
    } else if (s2 == out.csite) {
      equal = true;
      for (int k = 0; k < n; ++k) {
        if (c2[k] < out.ccfg[k]) { better = true; equal = false; break; }
        if (c2[k] > out.ccfg[k]) { equal = false; break; }
      }
    }
    if (better) {
      out.csite = s2;
      for (int k = 0; k < n; ++k) out.ccfg[k] = c2[k];
      ntie = 1;
      out.gbest = op;
    } else if (equal) {
      ++ntie;
    }
  }
  out.osize = nop / ntie;
}

inline int get_state(const int* cfg, int n, int r, int nsm) {
  for (int i = 0; i < n; ++i)
    if (cfg[i] / nsm == r) return cfg[i] % nsm;
  return 0;
}

// set_state (src/asr_augcls.f90:217-235): drop site r, then insert r*nsm + k (k > 0) keeping the codes sorted.
inline void set_state(const std::vector<int>& cfg, int r, int k, int nsm, std::vector<int>& out) {
  out.clear();
  for (int c : cfg)
    if (c / nsm != r) out.push_back(c);
  if (k > 0) {
    out.push_back(r * nsm + k);
    int i = (int)out.size() - 1;
    while (i > 0 && out[i - 1] > out[i]) { std::swap(out[i - 1], out[i]); --i; }
  }
}

}  // namespace

struct asr_augcls_tables {
  int nn = 0, nslot = 0, nop = 0, nsm = 0, nstab = 0;
  std::vector<int> hopr, hopg, flipr, flipg, lst, site, level, orb, stab_ptr, stab;
  // the node store and the group, kept for the target lookups (tnode / tgop)
  NodeStore st;
  int nsite = 0, start = 0;
  std::vector<int> perm, nst;
};

extern "C" int asr_host_build_augcls(int nsite, const int* nbr_ptr, const int* nbr, const int* cls, int nop, const int* perm,
                                     int start, int maxlevel, int maxrank, const unsigned char* flippable, const int* nsc,
                                     int ncls, int nthreads, asr_augcls_tables** out, char* err, int errlen) {
  auto fail = [&](const char* msg) {
    if (err && errlen > 0) { std::snprintf(err, errlen, "%s", msg); }
    return 1;
  };
  *out = nullptr;
  if (nsite <= 0 || nop <= 0) return fail("node-set builder: the site graph or the group is empty");
  for (int i = 0; i < nsite; ++i)
    if (perm[i] != i) return fail("node-set builder: the first group operation is not the identity permutation");
  for (int g = 0; g < nop; ++g) {
    const int* p = perm + (size_t)g * nsite;
    if (p[start] != start) return fail("node-set builder: a group operation moves the start site");
    for (int i = 0; i < nsite; ++i)
      if (cls[p[i]] != cls[i]) return fail("node-set builder: a group operation maps a site onto one of another class");
  }
  int nsm = 0;
  for (int c = 0; c < ncls; ++c) nsm = std::max(nsm, nsc[c]);
  const int rkmax = maxrank > 0 ? maxrank : 0x7fffffff;
  std::vector<int> nst(nsite);
  for (int r = 0; r < nsite; ++r) nst[r] = (flippable == nullptr || flippable[r]) ? nsc[cls[r]] : 1;
  int nslot = 0;
  for (int r = 0; r < nsite; ++r) nslot = std::max(nslot, nbr_ptr[r + 1] - nbr_ptr[r]);

  if (nthreads > 0) omp_set_num_threads(nthreads);
  NodeStore st;
  st.init(4096);
  {
    Canon cn;
    std::vector<int> c2;
    canon_node(perm, nsite, nop, nsm, start, nullptr, 0, cn, c2);
    if (cn.osize != 1) return fail("node-set builder: the start node has an orbit of more than one node under the group");
    bool isnew;
    int idx = st.find_or_add(cn.csite, cn.ccfg.data(), 0, 0, &isnew);
    st.orb[idx] = 1;
  }
  std::string errmsg;
  const int chunk = 2048;
  std::vector<std::vector<int>> cb(chunk);
  int i = 0;
  while (i < st.n) {
    int i2 = std::min(st.n - 1, i + chunk - 1);
    bool edge = false;
#pragma omp parallel
    {
      Canon cn;
      std::vector<int> c2, c0, cc, tmp;
#pragma omp for schedule(dynamic, 4)
      for (int ii = i; ii <= i2; ++ii) {
        std::vector<int>& rec = cb[ii - i];
        rec.clear();
        int lev = st.level[ii];
        if (lev >= maxlevel) continue;
        int r = st.site[ii];
        c0.assign(st.data.begin() + st.start[ii], st.data.begin() + st.start[ii] + st.len[ii]);
        int deg_r = nbr_ptr[r + 1] - nbr_ptr[r];
        for (int k1 = 0; k1 <= deg_r; ++k1) {
          int r1 = (k1 == 0) ? r : nbr[nbr_ptr[r] + k1 - 1];
          if (r1 == -1) continue;
          if (r1 < 0) { edge = true; continue; }
          int deg_1 = nbr_ptr[r1 + 1] - nbr_ptr[r1];
          for (int k2 = 0; k2 <= deg_1; ++k2) {
            int r2 = (k2 == 0) ? r1 : nbr[nbr_ptr[r1] + k2 - 1];
            if (r2 == -1) continue;
            if (r2 < 0) { edge = true; continue; }
            int sites[3], nsit = 1;
            sites[0] = r;
            if (r1 != r) sites[nsit++] = r1;
            if (r2 != r && r2 != r1) sites[nsit++] = r2;
            int ncomb = 1;
            for (int j = 0; j < nsit; ++j) ncomb *= nst[sites[j]];
            for (int f = 0; f < ncomb; ++f) {
              cc = c0;
              int q = f;
              for (int j = 0; j < nsit; ++j) {
                int ns = nst[sites[j]];
                if (ns == 1) continue;
                int dgt = q % ns;
                q /= ns;
                set_state(cc, sites[j], dgt, nsm, tmp);
                cc.swap(tmp);
              }
              if ((int)cc.size() > rkmax) continue;
              canon_node(perm, nsite, nop, nsm, r2, cc.data(), (int)cc.size(), cn, c2);
              rec.push_back(cn.csite);
              rec.push_back(cn.osize);
              rec.push_back((int)cn.ccfg.size());
              rec.insert(rec.end(), cn.ccfg.begin(), cn.ccfg.end());
            }
          }
        }
      }
    }
    if (edge) return fail("node-set builder: a path of the closure leaves the site graph (a neighbour index is missing)");
    for (int ii = i; ii <= i2; ++ii) {  // serial insertion, in the order of the serial loop
      int lev = st.level[ii];
      const std::vector<int>& rec = cb[ii - i];
      size_t k = 0;
      while (k < rec.size()) {
        int cs = rec[k], osz = rec[k + 1], nc = rec[k + 2];
        bool isnew;
        int idx = st.find_or_add(cs, rec.data() + k + 3, nc, lev + 1, &isnew);
        if (isnew) st.orb[idx] = osz;
        k += 3 + nc;
      }
    }
    i = i2 + 1;
  }

  auto* T = new asr_augcls_tables();
  const int nn = st.n;
  T->nn = nn; T->nslot = nslot; T->nop = nop; T->nsm = nsm;
  T->hopr.assign((size_t)nn * nslot, -1); T->hopg.assign((size_t)nn * nslot, 0);
  T->flipr.assign((size_t)nn * nsm, -1); T->flipg.assign((size_t)nn * nsm, 0);
  T->lst.assign(nn, 0); T->site.assign(nn, 0); T->level.assign(nn, 0); T->orb.assign(nn, 1);
  T->stab_ptr.assign(nn + 1, 0);
  std::vector<int> nstb(nn, 0);
  bool bad_stab = false;
#pragma omp parallel
  {
    std::vector<int> c2;
#pragma omp for schedule(dynamic, 64)
    for (int n = 0; n < nn; ++n) {
      int r = st.site[n];
      const int* c0 = &st.data[st.start[n]];
      int nc = st.len[n];
      int cnt = 0;
      c2.resize(nc);
      for (int g = 0; g < nop; ++g) {
        const int* p = perm + (size_t)g * nsite;
        if (p[r] != r) continue;
        for (int k = 0; k < nc; ++k) c2[k] = p[c0[k] / nsm] * nsm + c0[k] % nsm;
        sort_small(c2.data(), nc);
        bool eq = true;
        for (int k = 0; k < nc; ++k)
          if (c2[k] != c0[k]) { eq = false; break; }
        if (eq) ++cnt;
      }
      nstb[n] = cnt;
      if (cnt * st.orb[n] != nop) bad_stab = true;
    }
  }
  if (bad_stab) { delete T; return fail("node-set builder: a stabiliser does not have group order / orbit size operations"); }
  for (int n = 0; n < nn; ++n) T->stab_ptr[n + 1] = T->stab_ptr[n] + nstb[n];
  T->nstab = T->stab_ptr[nn];
  T->stab.assign(T->nstab, 0);
#pragma omp parallel
  {
    Canon cn;
    std::vector<int> c2, c0v, tmp;
#pragma omp for schedule(dynamic, 64)
    for (int n = 0; n < nn; ++n) {
      int r = st.site[n];
      const int* c0 = &st.data[st.start[n]];
      int nc = st.len[n];
      c0v.assign(c0, c0 + nc);
      int j = T->stab_ptr[n];
      c2.resize(nc);
      for (int g = 0; g < nop; ++g) {
        const int* p = perm + (size_t)g * nsite;
        if (p[r] != r) continue;
        for (int k = 0; k < nc; ++k) c2[k] = p[c0[k] / nsm] * nsm + c0[k] % nsm;
        sort_small(c2.data(), nc);
        bool eq = true;
        for (int k = 0; k < nc; ++k)
          if (c2[k] != c0[k]) { eq = false; break; }
        if (eq) T->stab[j++] = g;
      }
      T->level[n] = st.level[n];
      T->orb[n] = st.orb[n];
      T->site[n] = r;
      int l0 = get_state(c0, nc, r, nsm);
      T->lst[n] = l0;
      int ns_r = nsc[cls[r]];
      bool canflip = (flippable == nullptr || flippable[r]);
      for (int k = 0; k < ns_r; ++k) {
        if (k == l0) {
          T->flipr[(size_t)n * nsm + k] = n;
          T->flipg[(size_t)n * nsm + k] = 0;
        } else if (k == 0 || canflip) {
          set_state(c0v, r, k, nsm, tmp);
          canon_node(perm, nsite, nop, nsm, r, tmp.data(), (int)tmp.size(), cn, c2);
          int f = st.find(cn.csite, cn.ccfg.data(), (int)cn.ccfg.size());
          if (f >= 0) {
            T->flipr[(size_t)n * nsm + k] = f;
            T->flipg[(size_t)n * nsm + k] = cn.gbest;
          }
        }
      }
      int deg = nbr_ptr[r + 1] - nbr_ptr[r];
      for (int k = 0; k < deg; ++k) {
        int nb = nbr[nbr_ptr[r] + k];
        if (nb < 0) continue;  // no neighbour on the slot, or beyond the cluster: hopr stays -1
        canon_node(perm, nsite, nop, nsm, nb, c0, nc, cn, c2);
        int f = st.find(cn.csite, cn.ccfg.data(), (int)cn.ccfg.size());
        if (f >= 0) {
          T->hopr[(size_t)n * nslot + k] = f;
          T->hopg[(size_t)n * nslot + k] = cn.gbest;
        }
      }
    }
  }
  T->nsite = nsite;
  T->start = start;
  T->perm.assign(perm, perm + (size_t)nop * nsite);
  T->nst = nst;
  T->st = std::move(st);
  *out = T;
  return 0;
}

// Target nodes of build_augcls_operator (tsite / tnode / tgop, src/asr_augcls.f90:805-819): for target site tsite[t] and
// local state k = 0 .. nsm - 1 (the site alone in state k, every other site in its reference state) tnode[t * nsm + k] is the
// representative node and tgop[t * nsm + k] the operation that maps the node onto it; -1 when the node is not in the node
// set, the state does not exist on the site, or the site cannot flip (the start site, a non-flippable site).
extern "C" int asr_host_tables_targets(const asr_augcls_tables* T, int nt, const int* tsite, int* tnode, int* tgop) {
  if (!T) return 1;
  Canon cn;
  std::vector<int> c2, cfg, tmp;
  for (int t = 0; t < nt; ++t) {
    const int r = tsite[t];
    if (r < 0 || r >= T->nsite) return 1;
    for (int k = 0; k < T->nsm; ++k) {
      tnode[(size_t)t * T->nsm + k] = -1;
      tgop[(size_t)t * T->nsm + k] = 0;
      if (k > 0 && (k >= T->nst[r] || r == T->start)) continue;  // nst = 1 on a site that cannot flip
      cfg.clear();
      set_state(cfg, r, k, T->nsm, tmp);
      canon_node(T->perm.data(), T->nsite, T->nop, T->nsm, r, tmp.data(), (int)tmp.size(), cn, c2);
      const int f = T->st.find(cn.csite, cn.ccfg.data(), (int)cn.ccfg.size());
      if (f >= 0) {
        tnode[(size_t)t * T->nsm + k] = f;
        tgop[(size_t)t * T->nsm + k] = cn.gbest;
      }
    }
  }
  return 0;
}

extern "C" int asr_host_tables_dims(const asr_augcls_tables* T, int* dims) {
  if (!T) return 1;
  dims[0] = T->nn; dims[1] = T->nslot; dims[2] = T->nop; dims[3] = T->nsm; dims[4] = T->nstab;
  return 0;
}

extern "C" int asr_host_tables_get(const asr_augcls_tables* T, int* hopr, int* hopg, int* flipr, int* flipg, int* lst, int* site,
                                   int* level, int* orb, int* stab_ptr, int* stab) {
  if (!T) return 1;
  auto cp = [](const std::vector<int>& v, int* dst) { if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(int)); };
  cp(T->hopr, hopr); cp(T->hopg, hopg); cp(T->flipr, flipr); cp(T->flipg, flipg); cp(T->lst, lst); cp(T->site, site);
  cp(T->level, level); cp(T->orb, orb); cp(T->stab_ptr, stab_ptr); cp(T->stab, stab);
  return 0;
}

extern "C" void asr_host_tables_free(asr_augcls_tables* T) { delete T; }
