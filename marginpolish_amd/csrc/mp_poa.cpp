// mp_poa.cpp -- POA graph augmentation from the posterior lists (host side).
//
// The reference builds the graph read by read (poa_realign,
// impl/polisher.c:656-703): match weights into the nodes, then the read's
// complete inserts and complete deletes, each moved to its leftmost
// equivalent position and merged into the node's list.  Here the per-read
// work that does not touch the graph (ordering the gap lists, finding the
// complete runs, left-shifting them) is done for all reads in parallel, and
// each graph then applies its reads' events in read order, so the lists come
// out in the reference's order and the weight sums in the reference's
// summation order.
#include <algorithm>
#include <atomic>
#include <cctype>
#include <cmath>
#include <limits>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "mp_internal.hpp"
#include "mp_poa.h"

namespace {

constexpr int kAlphabet = 5;  // A C G T N (alphabet_constructNucleotide, stateMachine.c:63-71)
constexpr double kUintMax = 4294967295.0;  // the reference's initial run minimum (UINT_MAX)

inline int base_symbol(char c) {  // convertNucleotideCharToSymbol (stateMachine.c:25-42)
    switch (c) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 4;
    }
}

// A run-length string view: runs [0, n) of (b, c).
struct Str {
    const char *b;
    const uint64_t *c;
    int64_t n;
};

struct InsEntry {
    uint64_t hash;
    int64_t node;
    int64_t sym, len;  // placed string in the graph's symbol pool
    double wF, wR;
    int32_t next;      // next entry of the same node (-1: last)
};

inline uint64_t slot_key(uint64_t hash, int64_t node) { return hash ^ ((uint64_t)node * 0x9E3779B97F4A7C15ULL); }

struct DelEntry {
    int64_t len;
    double wF, wR;
    int32_t next;
};

// A complete insert of one read, placed: node `pos`, the read's runs
// [start, start + len) rotated right by `rot`, run minimum weight `w`.
struct InsEvent {
    int64_t pos, start, len, rot;
    double w;
    uint64_t hash;
};

struct DelEvent {
    int64_t pos, len, off;
    double w;
};

inline uint64_t mix(uint64_t h, char b, uint64_t c) {
    h ^= (uint64_t)(unsigned char)b;
    h *= 0x100000001b3ULL;
    h ^= c;
    h *= 0x100000001b3ULL;
    return h;
}

}  // namespace

struct mp_poa {
    int64_t maxRC = 0;
    std::vector<char> refB;
    std::vector<uint64_t> refC;
    std::vector<char> nodeBase;
    std::vector<uint64_t> nodeRC;
    std::vector<double> baseW, rcW;  // [node][5], [node][maxRC]
    // node observations: the match tuples of every augment call in order, and
    // per call (readNo, index of its first tuple); node = x + 1, offset = y
    std::vector<mp_aligned_pair> obsPairs;
    std::vector<std::pair<int64_t, int64_t>> obsSeg;
    std::vector<int32_t> insHead, insTail, delHead, delTail;
    std::vector<InsEntry> ins;
    std::vector<int32_t> insTab;  // open-addressing index of ins by (node, string); -1 = empty
    std::vector<char> insB;
    std::vector<uint64_t> insC;
    std::vector<DelEntry> del;
    // insert / delete observations in append order, with the entry each belongs to
    std::vector<int32_t> insObsEnt, delObsEnt;
    std::vector<mp_poa_observation> insObs, delObs;
    // last consensus (a size query followed by the fill does not recompute it);
    // `gen` counts changes to the graph
    std::atomic<uint64_t> gen{0};
    mutable uint64_t consGen = ~0ull;
    mutable int32_t consRle = -1;
    mutable double consPenalty = 0.0;
    mutable std::string consString;
    mutable std::vector<int64_t> consMap;

    Str ref() const { return Str{refB.data(), refC.data(), (int64_t)refB.size()}; }
    int64_t nodes() const { return (int64_t)nodeBase.size(); }
};

namespace {

// matchesReferenceSubstring (polisher.c:237-249): s[0, len) == ref[at, at + len).
inline bool ref_matches(const Str &ref, int64_t at, const char *sb, const uint64_t *sc, int64_t len, bool cmp) {
    for (int64_t l = 0; l < len; l++)
        if (ref.b[at + l] != sb[l] || (cmp && ref.c[at + l] != sc[l])) return false;
    return true;
}

// Smallest k such that s is whole copies of its first k runs (hasInternalRepeat, polisher.c:251-267).
int64_t min_period(const char *sb, const uint64_t *sc, int64_t n, bool cmp) {
    for (int64_t k = 1; k < n; k++) {
        if (n % k) continue;
        bool ok = true;
        for (int64_t i = k; i < n && ok; i++) ok = sb[i] == sb[i - k] && (!cmp || sc[i] == sc[i - k]);
        if (ok) return k;
    }
    return std::max<int64_t>(n, 1);
}

// getShift (polisher.c:269-297): the leftmost start reached by stepping back
// one period at a time while the reference repeats the period; then, for a
// one-run string with counts compared, one more run if the base agrees.
int64_t left_shift(const Str &ref, int64_t at, const char *sb, const uint64_t *sc, int64_t n, bool cmp) {
    const int64_t m = min_period(sb, sc, n, cmp);
    for (int64_t k = at - m; k >= 0 && ref_matches(ref, k, sb, sc, m, cmp); k -= m) at = k;
    if (n == 1 && cmp && at > 0 && ref.b[at - 1] == sb[0]) at--;
    return at;
}

// getMaxCommonSuffixLength (polisher.c:299-314) of ref[0, end) and s.
int64_t common_suffix(const Str &ref, int64_t end, const char *sb, const uint64_t *sc, int64_t n, bool cmp) {
    int64_t i = 0;
    while (end - i - 1 >= 0 && n - i - 1 >= 0) {
        if (ref.b[end - 1 - i] != sb[n - 1 - i] || (cmp && ref.c[end - 1 - i] != sc[n - 1 - i])) break;
        i++;
    }
    return i;
}

// Per-thread scratch for one read's events.
struct Scratch {
    std::vector<int32_t> cnt;        // counting-sort buckets
    std::vector<int32_t> mY;         // match y values bucketed by x
    std::vector<int32_t> mStart;     // bucket starts (x range + 1)
    int64_t mX0 = 0, mX1 = -1;
    std::vector<mp_aligned_pair> sorted;

    void index_matches(const mp_aligned_pair *m, int64_t n) {
        mX0 = INT64_MAX;
        mX1 = INT64_MIN;
        for (int64_t i = 0; i < n; i++) {
            mX0 = std::min<int64_t>(mX0, m[i].x);
            mX1 = std::max<int64_t>(mX1, m[i].x);
        }
        if (n == 0) {
            mX0 = 0;
            mX1 = -1;
        }
        const int64_t span = mX1 - mX0 + 1;
        mStart.assign(span + 1, 0);
        for (int64_t i = 0; i < n; i++) mStart[m[i].x - mX0 + 1]++;
        for (int64_t k = 0; k < span; k++) mStart[k + 1] += mStart[k];
        mY.resize(n);
        cnt.assign(mStart.begin(), mStart.end());
        for (int64_t i = 0; i < n; i++) mY[cnt[m[i].x - mX0]++] = m[i].y;
    }
    bool is_match(int64_t x, int64_t y) const {
        if (x < mX0 || x > mX1) return false;
        for (int32_t k = mStart[x - mX0]; k < mStart[x - mX0 + 1]; k++)
            if (mY[k] == y) return true;
        return false;
    }
    // Stable sort of gap tuples by (major, minor), major = x (inserts) or y (deletes).
    void sort_gaps(const mp_aligned_pair *g, int64_t n, bool byX) {
        sorted.resize(n);
        if (n == 0) return;
        int64_t lo = INT64_MAX, hi = INT64_MIN;
        for (int64_t i = 0; i < n; i++) {
            const int64_t k = byX ? g[i].x : g[i].y;
            lo = std::min(lo, k);
            hi = std::max(hi, k);
        }
        cnt.assign(hi - lo + 2, 0);
        for (int64_t i = 0; i < n; i++) cnt[(byX ? g[i].x : g[i].y) - lo + 1]++;
        for (int64_t k = 0; k + 1 < (int64_t)cnt.size(); k++) cnt[k + 1] += cnt[k];
        for (int64_t i = 0; i < n; i++) sorted[cnt[(byX ? g[i].x : g[i].y) - lo]++] = g[i];
        // bucket k now ends at cnt[k]; order each bucket by the minor coordinate (buckets are small)
        int64_t s = 0;
        for (int64_t k = 0; k + 1 < (int64_t)cnt.size(); k++) {
            const int64_t e = cnt[k];
            for (int64_t i = s + 1; i < e; i++) {
                const mp_aligned_pair t = sorted[i];
                const int32_t key = byX ? t.y : t.x;
                int64_t j = i;
                while (j > s && (byX ? sorted[j - 1].y : sorted[j - 1].x) > key) {
                    sorted[j] = sorted[j - 1];
                    j--;
                }
                sorted[j] = t;
            }
            s = e;
        }
    }
};

bool check_read(const Str &ref, const Str &read, const mp_aligned_pair *m, int64_t nm, const mp_aligned_pair *ins,
                int64_t ni, const mp_aligned_pair *del, int64_t nd) {
    for (int64_t i = 0; i < nm; i++)
        if (m[i].x < 0 || m[i].x >= ref.n || m[i].y < 0 || m[i].y >= read.n) return false;
    for (int64_t i = 0; i < ni; i++)
        if (ins[i].x < -1 || ins[i].x >= ref.n || ins[i].y < 0 || ins[i].y >= read.n) return false;
    for (int64_t i = 0; i < nd; i++)
        if (del[i].x < 0 || del[i].x >= ref.n || del[i].y < -1 || del[i].y >= read.n) return false;
    return true;
}

// The complete inserts and deletes of one read (polisher.c:341-530), placed
// and in the reference's enumeration order.
void read_events(const Str &ref, const Str &read, const mp_aligned_pair *m, int64_t nm, const mp_aligned_pair *ins,
                 int64_t ni, const mp_aligned_pair *del, int64_t nd, bool cmp, Scratch &S,
                 std::vector<InsEvent> &iev, std::vector<DelEvent> &dev) {
    iev.clear();
    dev.clear();
    S.index_matches(m, nm);
    // inserts: runs of equal x and consecutive y; sub-runs [k, l] flanked by
    // a match (x, y_k - 1) or the read start and a match (x + 1, y_l + 1) or the read end
    S.sort_gaps(ins, ni, true);
    for (int64_t i = 0; i < ni;) {
        const int64_t x0 = S.sorted[i].x, y0 = S.sorted[i].y;
        int64_t j = i + 1;
        while (j < ni && S.sorted[j].x == x0 && S.sorted[j].y == y0 + (j - i)) j++;
        for (int64_t k = i; k < j; k++) {
            const int64_t yp = y0 + (k - i) - 1;
            if (yp > -1 && !S.is_match(x0, yp)) continue;
            double w = kUintMax;
            for (int64_t l = k; l < j; l++) {
                w = std::min(w, (double)S.sorted[l].weight);
                const int64_t yn = y0 + (l - i) + 1;
                if (yn < read.n && !S.is_match(x0 + 1, yn)) continue;
                const int64_t start = y0 + (k - i), len = l - k + 1;
                const char *sb = read.b + start;
                const uint64_t *sc = read.c + start;
                int64_t pos = left_shift(ref, x0 + 1, sb, sc, len, cmp);
                const int64_t cs = common_suffix(ref, pos, sb, sc, len, cmp);
                const int64_t rot = cs % len;  // rotating by the whole length is the identity
                pos -= cs;
                uint64_t h = 0xcbf29ce484222325ULL ^ (uint64_t)len;
                for (int64_t q = 0; q < len; q++) {
                    const int64_t src = (q - rot + len) % len;
                    h = mix(h, sb[src], sc[src]);
                }
                iev.push_back(InsEvent{pos, start, len, rot, w, h});
            }
        }
        i = j;
    }
    // deletes: runs of equal y and consecutive x, flanked by a match
    // (x_k - 1, y) or the reference start and (x_l + 1, y + 1) or its end
    S.sort_gaps(del, nd, false);
    for (int64_t i = 0; i < nd;) {
        const int64_t x0 = S.sorted[i].x, y0 = S.sorted[i].y;
        int64_t j = i + 1;
        while (j < nd && S.sorted[j].y == y0 && S.sorted[j].x == x0 + (j - i)) j++;
        for (int64_t k = i; k < j; k++) {
            const int64_t xp = x0 + (k - i) - 1;
            if (xp > -1 && !S.is_match(xp, y0)) continue;
            double w = kUintMax;
            for (int64_t l = k; l < j; l++) {
                w = std::min(w, (double)S.sorted[l].weight);
                const int64_t xn = x0 + (l - i) + 1;
                if (xn < ref.n && !S.is_match(xn, y0 + 1)) continue;
                const int64_t found = x0 + (k - i), len = l - k + 1;
                int64_t pos = left_shift(ref, found, ref.b + found, ref.c + found, len, cmp);
                pos -= common_suffix(ref, pos, ref.b + found, ref.c + found, len, cmp);
                dev.push_back(DelEvent{pos, len, y0, w});
            }
        }
        i = j;
    }
}

// Graph updates of one read, in the reference's order: matches, inserts,
// deletes.  The three touch disjoint parts of the graph, so for a sequence
// of reads each part may be brought up to date on its own.
void apply_matches(mp_poa *g, const Str &read, int64_t readNo, const mp_aligned_pair *m, int64_t nm) {
    g->gen.fetch_add(1, std::memory_order_relaxed);
    const int64_t mrc = g->maxRC;
    if (nm > 0) {
        g->obsSeg.emplace_back(readNo, (int64_t)g->obsPairs.size());
        g->obsPairs.insert(g->obsPairs.end(), m, m + nm);
    }
    double *bw = g->baseW.data(), *rw = g->rcW.data();
    for (int64_t i = 0; i < nm; i++) {
        const int64_t node = (int64_t)m[i].x + 1, y = m[i].y;
        const double w = (double)m[i].weight;
        bw[node * kAlphabet + base_symbol(read.b[y])] += w;
        const uint64_t rc = read.c[y] >= (uint64_t)mrc ? (uint64_t)(mrc - 1) : read.c[y];
        rw[node * mrc + (int64_t)rc] += w;
    }
}

void apply_inserts(mp_poa *g, const Str &read, bool strand, int64_t readNo, const std::vector<InsEvent> &iev) {
    g->gen.fetch_add(1, std::memory_order_relaxed);
    for (const InsEvent &e : iev) {
        const char *sb = read.b + e.start;
        const uint64_t *sc = read.c + e.start;
        // the node's entry with the same run-length string, through the (node, string) index
        if (2 * (g->ins.size() + 1) > g->insTab.size()) {
            const size_t cap = std::max<size_t>(1024, 2 * g->insTab.size());
            g->insTab.assign(cap, -1);
            for (size_t q = 0; q < g->ins.size(); q++) {
                size_t h = slot_key(g->ins[q].hash, g->ins[q].node) & (cap - 1);
                while (g->insTab[h] >= 0) h = (h + 1) & (cap - 1);
                g->insTab[h] = (int32_t)q;
            }
        }
        const size_t mask = g->insTab.size() - 1;
        size_t slot = slot_key(e.hash, e.pos) & mask;
        int32_t hit = -1;
        for (; g->insTab[slot] >= 0; slot = (slot + 1) & mask) {
            const InsEntry &E = g->ins[g->insTab[slot]];
            if (E.hash != e.hash || E.node != e.pos || E.len != e.len) continue;
            bool same = true;
            for (int64_t t = 0; t < e.len && same; t++) {
                const int64_t src = (t - e.rot + e.len) % e.len;
                same = g->insB[E.sym + t] == sb[src] && g->insC[E.sym + t] == sc[src];
            }
            if (same) {
                hit = g->insTab[slot];
                break;
            }
        }
        if (hit < 0) {
            hit = (int32_t)g->ins.size();
            g->insTab[slot] = hit;
            const InsEntry E{e.hash, e.pos, (int64_t)g->insB.size(), e.len, 0.0, 0.0, -1};
            for (int64_t t = 0; t < e.len; t++) {
                const int64_t src = (t - e.rot + e.len) % e.len;
                g->insB.push_back(sb[src]);
                g->insC.push_back(sc[src]);
            }
            g->ins.push_back(E);
            if (g->insTail[e.pos] >= 0)
                g->ins[g->insTail[e.pos]].next = hit;
            else
                g->insHead[e.pos] = hit;
            g->insTail[e.pos] = hit;
        }
        InsEntry &E = g->ins[hit];
        (strand ? E.wF : E.wR) += e.w;
        g->insObsEnt.push_back(hit);
        g->insObs.push_back(mp_poa_observation{readNo, e.start, e.w});
    }
}

// One node range's share of a graph's insert task.  Entries the graph held
// before this call are updated in place (their nodes belong to this range
// only); entries new in this call are built here and merged into the graph
// afterwards, in this range's creation order, so every node's list and every
// entry's sums and observations come out as the sequential loop makes them.
struct InsShard {
    int64_t n0 = 0, n1 = 0;
    std::vector<InsEntry> ins;
    std::vector<char> b;
    std::vector<uint64_t> c;
    std::vector<int32_t> tab;
    std::vector<int32_t> obsEnt;  // >= 0: local entry; < 0: the graph's entry -(value + 1)
    std::vector<mp_poa_observation> obs;
};

inline bool same_insert(const char *eb, const uint64_t *ec, const InsEntry &E, const InsEvent &e, const char *sb,
                        const uint64_t *sc) {
    if (E.hash != e.hash || E.node != e.pos || E.len != e.len) return false;
    for (int64_t t = 0; t < e.len; t++) {
        const int64_t src = (t - e.rot + e.len) % e.len;
        if (eb[E.sym + t] != sb[src] || ec[E.sym + t] != sc[src]) return false;
    }
    return true;
}

void shard_inserts(mp_poa *g, InsShard &S, const Str &read, bool strand, int64_t readNo,
                   const std::vector<InsEvent> &iev) {
    for (const InsEvent &e : iev) {
        if (e.pos < S.n0 || e.pos >= S.n1) continue;
        const char *sb = read.b + e.start;
        const uint64_t *sc = read.c + e.start;
        int32_t hit = -1;
        if (!g->insTab.empty()) {  // the graph's own entries (index read-only here)
            const size_t mask = g->insTab.size() - 1;
            for (size_t slot = slot_key(e.hash, e.pos) & mask; g->insTab[slot] >= 0; slot = (slot + 1) & mask)
                if (same_insert(g->insB.data(), g->insC.data(), g->ins[g->insTab[slot]], e, sb, sc)) {
                    hit = g->insTab[slot];
                    break;
                }
        }
        if (hit >= 0) {
            InsEntry &E = g->ins[hit];
            (strand ? E.wF : E.wR) += e.w;
            S.obsEnt.push_back(-(hit + 1));
            S.obs.push_back(mp_poa_observation{readNo, e.start, e.w});
            continue;
        }
        if (2 * (S.ins.size() + 1) > S.tab.size()) {
            const size_t cap = std::max<size_t>(256, 2 * S.tab.size());
            S.tab.assign(cap, -1);
            for (size_t q = 0; q < S.ins.size(); q++) {
                size_t h = slot_key(S.ins[q].hash, S.ins[q].node) & (cap - 1);
                while (S.tab[h] >= 0) h = (h + 1) & (cap - 1);
                S.tab[h] = (int32_t)q;
            }
        }
        const size_t mask = S.tab.size() - 1;
        size_t slot = slot_key(e.hash, e.pos) & mask;
        for (; S.tab[slot] >= 0; slot = (slot + 1) & mask)
            if (same_insert(S.b.data(), S.c.data(), S.ins[S.tab[slot]], e, sb, sc)) {
                hit = S.tab[slot];
                break;
            }
        if (hit < 0) {
            hit = (int32_t)S.ins.size();
            S.tab[slot] = hit;
            S.ins.push_back(InsEntry{e.hash, e.pos, (int64_t)S.b.size(), e.len, 0.0, 0.0, -1});
            for (int64_t t = 0; t < e.len; t++) {
                const int64_t src = (t - e.rot + e.len) % e.len;
                S.b.push_back(sb[src]);
                S.c.push_back(sc[src]);
            }
        }
        InsEntry &E = S.ins[hit];
        (strand ? E.wF : E.wR) += e.w;
        S.obsEnt.push_back(hit);
        S.obs.push_back(mp_poa_observation{readNo, e.start, e.w});
    }
}

// The shards' new entries into the graph (shards in node order), then the index.
void merge_shards(mp_poa *g, std::vector<InsShard> &shards) {
    g->gen.fetch_add(1, std::memory_order_relaxed);
    for (InsShard &S : shards) {
        const int32_t base = (int32_t)g->ins.size();
        const int64_t symBase = (int64_t)g->insB.size();
        g->insB.insert(g->insB.end(), S.b.begin(), S.b.end());
        g->insC.insert(g->insC.end(), S.c.begin(), S.c.end());
        for (const InsEntry &E : S.ins) {
            const int32_t q = (int32_t)g->ins.size();
            InsEntry F = E;
            F.sym += symBase;
            F.next = -1;
            g->ins.push_back(F);
            if (g->insTail[E.node] >= 0)
                g->ins[g->insTail[E.node]].next = q;
            else
                g->insHead[E.node] = q;
            g->insTail[E.node] = q;
        }
        for (size_t i = 0; i < S.obsEnt.size(); i++) {
            g->insObsEnt.push_back(S.obsEnt[i] >= 0 ? base + S.obsEnt[i] : -S.obsEnt[i] - 1);
            g->insObs.push_back(S.obs[i]);
        }
        std::vector<char>().swap(S.b);
    }
    size_t cap = 1024;
    while (cap < 2 * (g->ins.size() + 1)) cap *= 2;
    g->insTab.assign(cap, -1);
    for (size_t q = 0; q < g->ins.size(); q++) {
        size_t h = slot_key(g->ins[q].hash, g->ins[q].node) & (cap - 1);
        while (g->insTab[h] >= 0) h = (h + 1) & (cap - 1);
        g->insTab[h] = (int32_t)q;
    }
}

void apply_deletes(mp_poa *g, bool strand, int64_t readNo, const std::vector<DelEvent> &dev) {
    g->gen.fetch_add(1, std::memory_order_relaxed);
    for (const DelEvent &e : dev) {
        int32_t hit = -1;
        for (int32_t q = g->delHead[e.pos]; q >= 0; q = g->del[q].next)
            if (g->del[q].len == e.len) {
                hit = q;
                break;
            }
        if (hit < 0) {
            hit = (int32_t)g->del.size();
            g->del.push_back(DelEntry{e.len, 0.0, 0.0, -1});
            if (g->delTail[e.pos] >= 0)
                g->del[g->delTail[e.pos]].next = hit;
            else
                g->delHead[e.pos] = hit;
            g->delTail[e.pos] = hit;
        }
        DelEntry &E = g->del[hit];
        (strand ? E.wF : E.wR) += e.w;
        g->delObsEnt.push_back(hit);
        g->delObs.push_back(mp_poa_observation{readNo, e.off, e.w});
    }
}

bool valid_string(const mp_rle_string *s) {
    return s && s->length >= 0 && (s->length == 0 || (s->bases && s->repeatCounts));
}

// Entries of one kind in read-out order: nodes ascending, each node's list in order.
template <class E>
void list_order(const std::vector<int32_t> &head, const std::vector<E> &ent, std::vector<std::pair<int64_t, int32_t>> &out) {
    out.clear();
    for (int64_t n = 0; n < (int64_t)head.size(); n++)
        for (int32_t q = head[n]; q >= 0; q = ent[q].next) out.emplace_back(n, q);
}

}  // namespace

namespace mp {

mp_status poa_augment_pairs(mp_poa *const *poas, int64_t nPoas, const int64_t *pairPoa, const int64_t *pairReadNo,
                            const mp_rle_string *reads, const uint8_t *strand, int64_t npairs,
                            const mp_aligned_pair *const lists[3], const int64_t *const starts[3], bool cmp,
                            int threads) {
    for (int64_t i = 0; i < npairs; i++) {
        if (pairPoa[i] < 0 || pairPoa[i] >= nPoas || !poas[pairPoa[i]] || !valid_string(&reads[i])) {
            set_error("mp_batch_poa_augment: pair " + std::to_string(i) + " has no graph or a bad read string");
            return MP_EINVAL;
        }
    }
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    threads = (int)std::max<int64_t>(1, std::min<int64_t>(threads, std::max<int64_t>(npairs, 1)));
    // each graph's pairs, in batch order
    std::vector<std::vector<int64_t>> byPoa(nPoas);
    for (int64_t i = 0; i < npairs; i++) byPoa[pairPoa[i]].push_back(i);
    std::vector<int64_t> poaOrder;
    for (int64_t p = 0; p < nPoas; p++)
        if (!byPoa[p].empty()) poaOrder.push_back(p);
    // larger graphs (more list tuples) first
    std::vector<int64_t> poaWork(nPoas, 0);
    for (int64_t i = 0; i < npairs; i++)
        for (int l = 0; l < 3; l++) poaWork[pairPoa[i]] += starts[l][i + 1] - starts[l][i];
    std::stable_sort(poaOrder.begin(), poaOrder.end(), [&](int64_t a, int64_t b) { return poaWork[a] > poaWork[b]; });

    std::vector<std::vector<InsEvent>> iev(npairs);
    std::vector<std::vector<DelEvent>> dev(npairs);
    auto pair_lists = [&](int64_t i, const mp_aligned_pair *&m, int64_t &nm, const mp_aligned_pair *&gx, int64_t &nx,
                          const mp_aligned_pair *&gy, int64_t &ny) {
        m = lists[0] + starts[0][i];
        gx = lists[1] + starts[1][i];
        gy = lists[2] + starts[2][i];
        nm = starts[0][i + 1] - starts[0][i];
        nx = starts[1][i + 1] - starts[1][i];
        ny = starts[2][i + 1] - starts[2][i];
    };
    // Work items, one pool of threads:
    //  - events of a read (graphs read-only), reads grouped by graph, larger graphs first;
    //  - a graph's insert task and delete task, ready once all its reads' events exist;
    //  - a graph's match task (no dependency).
    // Each task applies the graph's reads in batch order, so lists and sums come
    // out in the reference's order.
    std::vector<int64_t> evOrder;
    evOrder.reserve((size_t)npairs);
    for (int64_t p : poaOrder) evOrder.insert(evOrder.end(), byPoa[p].begin(), byPoa[p].end());
    std::unique_ptr<std::atomic<int64_t>[]> pending(new std::atomic<int64_t>[std::max<int64_t>(nPoas, 1)]);
    for (int64_t p = 0; p < nPoas; p++) pending[p].store((int64_t)byPoa[p].size());
    std::mutex readyMu;
    std::vector<std::pair<int64_t, int>> ready;
    size_t readyHead = 0;
    std::atomic<int64_t> nextEv{0}, nextMatch{0}, done{0};
    const int64_t totalTasks = 3 * (int64_t)poaOrder.size();
    // a graph's insert task with many events runs as node-range shards; the
    // last shard to finish merges them (kind 10 + shard index)
    struct ShardSet {
        std::vector<InsShard> s;
        std::atomic<int> left{0};
    };
    std::vector<std::unique_ptr<ShardSet>> shardSets((size_t)std::max<int64_t>(nPoas, 1));
    constexpr int64_t kEventsPerShard = 20000;
    constexpr int kMaxShards = 8;
    auto push_ready = [&](int64_t p, int kind) {
        std::lock_guard<std::mutex> lk(readyMu);
        ready.emplace_back(p, kind);
    };
    auto run_task = [&](int64_t p, int kind) {
        mp_poa *g = poas[p];
        if (kind == 1) {
            int64_t nev = 0;
            for (int64_t i : byPoa[p]) nev += (int64_t)iev[i].size();
            const int ns = (int)std::min<int64_t>(kMaxShards, nev / kEventsPerShard);
            if (ns >= 2) {
                shardSets[p].reset(new ShardSet);
                ShardSet &ss = *shardSets[p];
                ss.s.resize((size_t)ns);
                const int64_t n = g->nodes();
                for (int k = 0; k < ns; k++) {
                    ss.s[k].n0 = n * k / ns;
                    ss.s[k].n1 = n * (k + 1) / ns;
                }
                ss.left.store(ns);
                for (int k = 0; k < ns; k++) push_ready(p, 10 + k);
                return;  // counted done when the last shard merges
            }
        }
        if (kind >= 10) {
            ShardSet &ss = *shardSets[p];
            InsShard &S = ss.s[kind - 10];
            for (int64_t i : byPoa[p]) {
                const Str rd{reads[i].bases, reads[i].repeatCounts, reads[i].length};
                shard_inserts(g, S, rd, strand[i] != 0, pairReadNo ? pairReadNo[i] : i, iev[i]);
            }
            if (ss.left.fetch_sub(1) == 1) {
                merge_shards(g, ss.s);
                for (int64_t i : byPoa[p]) std::vector<InsEvent>().swap(iev[i]);
                shardSets[p].reset();
                done.fetch_add(1);
            }
            return;
        }
        if (kind == 0) {
            int64_t nmTot = 0;
            for (int64_t i : byPoa[p]) nmTot += starts[0][i + 1] - starts[0][i];
            g->obsPairs.reserve(g->obsPairs.size() + nmTot);
        }
        for (int64_t i : byPoa[p]) {
            const Str rd{reads[i].bases, reads[i].repeatCounts, reads[i].length};
            const int64_t rn = pairReadNo ? pairReadNo[i] : i;
            if (kind == 0) {
                apply_matches(g, rd, rn, lists[0] + starts[0][i], starts[0][i + 1] - starts[0][i]);
            } else if (kind == 1) {
                apply_inserts(g, rd, strand[i] != 0, rn, iev[i]);
                std::vector<InsEvent>().swap(iev[i]);
            } else {
                apply_deletes(g, strand[i] != 0, rn, dev[i]);
                std::vector<DelEvent>().swap(dev[i]);
            }
        }
        done.fetch_add(1);
    };
    ThreadPoolRun pool;
    auto worker = [&](int) {
        Scratch S;
        for (;;) {
            if (pool.failed()) break;  // another worker ran out of memory: stop, join, report
            std::pair<int64_t, int> task{-1, 0};
            {
                std::lock_guard<std::mutex> lk(readyMu);
                if (readyHead < ready.size()) task = ready[readyHead++];
            }
            if (task.first >= 0) {
                run_task(task.first, task.second);
                continue;
            }
            const int64_t q = nextEv.fetch_add(1);
            if (q < (int64_t)evOrder.size()) {
                const int64_t i = evOrder[q];
                const mp_poa *g = poas[pairPoa[i]];
                const Str rd{reads[i].bases, reads[i].repeatCounts, reads[i].length};
                const mp_aligned_pair *m, *gx, *gy;
                int64_t nm, nx, ny;
                pair_lists(i, m, nm, gx, nx, gy, ny);
                read_events(g->ref(), rd, m, nm, gy, ny, gx, nx, cmp, S, iev[i], dev[i]);
                if (pending[pairPoa[i]].fetch_sub(1) == 1) {
                    push_ready(pairPoa[i], 1);
                    push_ready(pairPoa[i], 2);
                }
                continue;
            }
            const int64_t mt = nextMatch.fetch_add(1);
            if (mt < (int64_t)poaOrder.size()) {
                run_task(poaOrder[mt], 0);
                continue;
            }
            if (done.load() >= totalTasks) break;
            std::this_thread::yield();
        }
    };
    try {
        // every pair's coordinates are checked before any graph changes
        std::atomic<int64_t> nextCheck{0}, badPair{INT64_MAX};
        auto checker = [&]() {
            for (int64_t i0; (i0 = nextCheck.fetch_add(64)) < npairs;)
                for (int64_t i = i0; i < std::min<int64_t>(npairs, i0 + 64); i++) {
                    const mp_aligned_pair *m, *gx, *gy;
                    int64_t nm, nx, ny;
                    pair_lists(i, m, nm, gx, nx, gy, ny);
                    const Str rd{reads[i].bases, reads[i].repeatCounts, reads[i].length};
                    if (!check_read(poas[pairPoa[i]]->ref(), rd, m, nm, gy, ny, gx, nx)) {
                        int64_t cur = badPair.load();
                        while (i < cur && !badPair.compare_exchange_weak(cur, i)) {
                        }
                    }
                }
        };
        parallel_run(threads, [&](int) { checker(); });
        if (badPair.load() != INT64_MAX) {
            set_error("mp_batch_poa_augment: pair " + std::to_string(badPair.load()) +
                      " has list coordinates outside its graph's reference or its read");
            return MP_EINVAL;
        }
        pool.run(threads, worker);
    } catch (const std::bad_alloc &) {
        set_error("mp_batch_poa_augment: out of host memory (the graphs may be partly augmented)");
        return MP_ENOMEM;
    }
    return MP_OK;
}

}  // namespace mp

extern "C" mp_status mp_poa_reference_graph(const mp_rle_string *reference, int64_t maxRepeatCount, mp_poa **out) {
    if (!out || !valid_string(reference) || maxRepeatCount < 1) {
        mp::set_error("mp_poa_reference_graph: bad argument");
        return MP_EINVAL;
    }
    *out = nullptr;
    try {
        std::unique_ptr<mp_poa> g(new mp_poa);
        const int64_t n = reference->length;
        g->maxRC = maxRepeatCount;
        g->refB.assign(reference->bases, reference->bases + n);
        g->refC.assign(reference->repeatCounts, reference->repeatCounts + n);
        g->nodeBase.resize(n + 1);
        g->nodeRC.resize(n + 1);
        g->nodeBase[0] = 'N';  // prefix node for events before the first reference run
        g->nodeRC[0] = 1;
        for (int64_t i = 0; i < n; i++) {  // poaNode_construct (polisher.c:90-107)
            const char u = (char)std::toupper((unsigned char)reference->bases[i]);
            g->nodeBase[i + 1] = base_symbol(u) < 4 ? u : 'N';
            g->nodeRC[i + 1] = reference->repeatCounts[i];
        }
        g->baseW.assign((n + 1) * kAlphabet, 0.0);
        g->rcW.assign((n + 1) * maxRepeatCount, 0.0);
        g->insHead.assign(n + 1, -1);
        g->insTail.assign(n + 1, -1);
        g->delHead.assign(n + 1, -1);
        g->delTail.assign(n + 1, -1);
        *out = g.release();
    } catch (const std::bad_alloc &) {
        mp::set_error("mp_poa_reference_graph: out of host memory");
        return MP_ENOMEM;
    }
    return MP_OK;
}

extern "C" void mp_poa_free(mp_poa *poa) { delete poa; }

extern "C" mp_status mp_poa_augment(mp_poa *poa, const mp_rle_string *read, int32_t readStrand, int64_t readNo,
                                    const mp_aligned_pair *matches, int64_t nMatches, const mp_aligned_pair *inserts,
                                    int64_t nInserts, const mp_aligned_pair *deletes, int64_t nDeletes,
                                    int32_t compareRepeatCounts) {
    if (!poa || !valid_string(read) || nMatches < 0 || nInserts < 0 || nDeletes < 0 || (nMatches && !matches) ||
        (nInserts && !inserts) || (nDeletes && !deletes)) {
        mp::set_error("mp_poa_augment: bad argument");
        return MP_EINVAL;
    }
    const Str rd{read->bases, read->repeatCounts, read->length};
    if (!check_read(poa->ref(), rd, matches, nMatches, inserts, nInserts, deletes, nDeletes)) {
        mp::set_error("mp_poa_augment: list coordinates outside the reference or the read");
        return MP_EINVAL;
    }
    try {
        Scratch S;
        std::vector<InsEvent> iev;
        std::vector<DelEvent> dev;
        read_events(poa->ref(), rd, matches, nMatches, inserts, nInserts, deletes, nDeletes, compareRepeatCounts != 0,
                    S, iev, dev);
        apply_matches(poa, rd, readNo, matches, nMatches);
        apply_inserts(poa, rd, readStrand != 0, readNo, iev);
        apply_deletes(poa, readStrand != 0, readNo, dev);
    } catch (const std::bad_alloc &) {
        mp::set_error("mp_poa_augment: out of host memory");
        return MP_ENOMEM;
    }
    return MP_OK;
}

extern "C" mp_status mp_poa_get_sizes(const mp_poa *poa, mp_poa_sizes *s) {
    if (!poa || !s) {
        mp::set_error("mp_poa_get_sizes: null argument");
        return MP_EINVAL;
    }
    s->nodes = poa->nodes();
    s->maxRepeatCount = poa->maxRC;
    s->inserts = (int64_t)poa->ins.size();
    s->insertSymbols = (int64_t)poa->insB.size();
    s->deletes = (int64_t)poa->del.size();
    s->nodeObservations = (int64_t)poa->obsPairs.size();
    s->insertObservations = (int64_t)poa->insObs.size();
    s->deleteObservations = (int64_t)poa->delObs.size();
    return MP_OK;
}

extern "C" mp_status mp_poa_get_nodes(const mp_poa *poa, char *bases, uint64_t *repeatCounts, double *baseWeights,
                                      double *repeatCountWeights) {
    if (!poa) {
        mp::set_error("mp_poa_get_nodes: null graph");
        return MP_EINVAL;
    }
    const size_t n = poa->nodeBase.size();
    if (bases) std::memcpy(bases, poa->nodeBase.data(), n);
    if (repeatCounts) std::memcpy(repeatCounts, poa->nodeRC.data(), n * sizeof(uint64_t));
    if (baseWeights) std::memcpy(baseWeights, poa->baseW.data(), poa->baseW.size() * sizeof(double));
    if (repeatCountWeights) std::memcpy(repeatCountWeights, poa->rcW.data(), poa->rcW.size() * sizeof(double));
    return MP_OK;
}

extern "C" mp_status mp_poa_get_inserts(const mp_poa *poa, int64_t *node, double *wF, double *wR,
                                        int64_t *symbolStart, int64_t *symbolLength, char *bases,
                                        uint64_t *repeatCounts) {
    if (!poa) {
        mp::set_error("mp_poa_get_inserts: null graph");
        return MP_EINVAL;
    }
    std::vector<std::pair<int64_t, int32_t>> ord;
    list_order(poa->insHead, poa->ins, ord);
    int64_t sym = 0;
    for (size_t k = 0; k < ord.size(); k++) {
        const InsEntry &e = poa->ins[ord[k].second];
        if (node) node[k] = ord[k].first;
        if (wF) wF[k] = e.wF;
        if (wR) wR[k] = e.wR;
        if (symbolStart) symbolStart[k] = sym;
        if (symbolLength) symbolLength[k] = e.len;
        if (bases) std::memcpy(bases + sym, poa->insB.data() + e.sym, (size_t)e.len);
        if (repeatCounts) std::memcpy(repeatCounts + sym, poa->insC.data() + e.sym, (size_t)e.len * sizeof(uint64_t));
        sym += e.len;
    }
    return MP_OK;
}

extern "C" mp_status mp_poa_get_deletes(const mp_poa *poa, int64_t *node, int64_t *length, double *wF, double *wR) {
    if (!poa) {
        mp::set_error("mp_poa_get_deletes: null graph");
        return MP_EINVAL;
    }
    std::vector<std::pair<int64_t, int32_t>> ord;
    list_order(poa->delHead, poa->del, ord);
    for (size_t k = 0; k < ord.size(); k++) {
        const DelEntry &e = poa->del[ord[k].second];
        if (node) node[k] = ord[k].first;
        if (length) length[k] = e.len;
        if (wF) wF[k] = e.wF;
        if (wR) wR[k] = e.wR;
    }
    return MP_OK;
}

extern "C" mp_status mp_poa_get_observations(const mp_poa *poa, int32_t kind, int64_t *start,
                                             mp_poa_observation *obs) {
    if (!poa || !start || kind < 0 || kind > 2) {
        mp::set_error("mp_poa_get_observations: bad argument");
        return MP_EINVAL;
    }
    if (kind == 0) {  // stable counting sort of the append-order observations by node
        const int64_t n = poa->nodes();
        const std::vector<mp_aligned_pair> &op = poa->obsPairs;
        std::vector<int64_t> cnt(n + 1, 0);
        for (const mp_aligned_pair &t : op) cnt[(int64_t)t.x + 1]++;
        int64_t run = 0;
        for (int64_t k = 0; k <= n; k++) {
            const int64_t c = cnt[k];
            cnt[k] = run;
            run += c;
        }
        std::memcpy(start, cnt.data(), (n + 1) * sizeof(int64_t));
        start[n] = run;
        if (obs)
            for (size_t sg = 0; sg < poa->obsSeg.size(); sg++) {
                const int64_t a = poa->obsSeg[sg].second;
                const int64_t e = sg + 1 < poa->obsSeg.size() ? poa->obsSeg[sg + 1].second : (int64_t)op.size();
                const int64_t readNo = poa->obsSeg[sg].first;
                for (int64_t i = a; i < e; i++)
                    obs[cnt[(int64_t)op[i].x + 1]++] = mp_poa_observation{readNo, op[i].y, (double)op[i].weight};
            }
        return MP_OK;
    }
    std::vector<std::pair<int64_t, int32_t>> ord;
    if (kind == 1)
        list_order(poa->insHead, poa->ins, ord);
    else
        list_order(poa->delHead, poa->del, ord);
    const std::vector<int32_t> &ent = kind == 1 ? poa->insObsEnt : poa->delObsEnt;
    const std::vector<mp_poa_observation> &ob = kind == 1 ? poa->insObs : poa->delObs;
    // entry -> read-out position, then a stable counting sort of the observations by it
    std::vector<int64_t> rank(ord.size()), cnt(ord.size() + 1, 0);
    for (size_t k = 0; k < ord.size(); k++) rank[ord[k].second] = (int64_t)k;
    for (int32_t q : ent) cnt[rank[q] + 1]++;
    for (size_t k = 0; k < ord.size(); k++) cnt[k + 1] += cnt[k];
    std::memcpy(start, cnt.data(), (ord.size() + 1) * sizeof(int64_t));
    if (obs)
        for (size_t i = 0; i < ob.size(); i++) obs[cnt[rank[ent[i]]]++] = ob[i];
    return MP_OK;
}

extern "C" mp_status mp_poa_estimate_repeat_counts(const mp_poa *poa, const mp_model *m, const mp_rle_string *reads,
                                                   const int32_t *readStrand, int64_t nReads, uint64_t *repeatCounts,
                                                   int32_t threads) {
    if (!poa || !m || !repeatCounts || nReads < 0 || (nReads > 0 && (!reads || !readStrand))) {
        mp::set_error("mp_poa_estimate_repeat_counts: bad argument");
        return MP_EINVAL;
    }
    constexpr int64_t kRL = mp::kMaxRepeat;  // MAXIMUM_REPEAT_LENGTH: the repeat table's bound (exclusive)
    if ((int64_t)m->repeatLog10.size() < 4 * kRL * kRL) {
        mp::set_error("mp_poa_estimate_repeat_counts: the model has no repeat-count table");
        return MP_EINVAL;
    }
    const int64_t n = poa->nodes();
    const std::vector<mp_aligned_pair> &op = poa->obsPairs;
    // every observation's read must exist and cover its offset
    for (size_t sg = 0; sg < poa->obsSeg.size(); sg++) {
        const int64_t rn = poa->obsSeg[sg].first;
        if (rn < 0 || rn >= nReads || !valid_string(&reads[rn])) {
            mp::set_error("mp_poa_estimate_repeat_counts: observation of read " + std::to_string(rn) +
                          " outside the read array");
            return MP_EINVAL;
        }
        const int64_t e = sg + 1 < poa->obsSeg.size() ? poa->obsSeg[sg + 1].second : (int64_t)op.size();
        for (int64_t i = poa->obsSeg[sg].second; i < e; i++)
            if (op[i].y < 0 || op[i].y >= reads[rn].length) {
                mp::set_error("mp_poa_estimate_repeat_counts: observation offset outside read " + std::to_string(rn));
                return MP_EINVAL;
            }
    }
    try {
        // each node's observations in append order: (observed run length, strand, weight)
        struct Obs {
            uint64_t observed;
            int32_t strand;
            double weight;
        };
        std::vector<int64_t> start(n + 1, 0);
        for (const mp_aligned_pair &t : op) start[(int64_t)t.x + 2]++;
        for (int64_t k = 0; k < n; k++) start[k + 1] += start[k];
        std::vector<Obs> ob(op.size());
        for (size_t sg = 0; sg < poa->obsSeg.size(); sg++) {
            const int64_t rn = poa->obsSeg[sg].first;
            const int64_t e = sg + 1 < poa->obsSeg.size() ? poa->obsSeg[sg + 1].second : (int64_t)op.size();
            for (int64_t i = poa->obsSeg[sg].second; i < e; i++)
                ob[start[(int64_t)op[i].x + 1]++] =
                    Obs{reads[rn].repeatCounts[op[i].y], readStrand[rn] ? 1 : 0, (double)op[i].weight};
        }
        // start[k] is now the end of node k's observations; node k's begin is start[k - 1] (0 for node 0)
        const double *tab = m->repeatLog10.data();
        auto node_count = [&](int64_t k) -> uint64_t {
            const int64_t a = k > 0 ? start[k - 1] : 0, e = start[k];
            int b = base_symbol(poa->nodeBase[k]);
            if (b == 4) b = 0;  // N reads the A row (polisher.c:1758)
            int64_t lo = kRL, hi = 0;
            for (int64_t i = a; i < e; i++) {
                lo = std::min<int64_t>(lo, (int64_t)std::min<uint64_t>(ob[i].observed, (uint64_t)INT64_MAX));
                hi = std::max<int64_t>(hi, (int64_t)std::min<uint64_t>(ob[i].observed, (uint64_t)INT64_MAX));
            }
            if (hi >= kRL) hi = kRL - 1;
            int64_t c;
            if (lo == kRL) {
                c = 0;  // no observation below the bound (otherwise lo <= hi)
            } else {
                double best = 0.0;
                c = lo;
                for (int64_t u = lo; u <= hi; u++) {
                    double lp = 0.0;
                    for (int64_t i = a; i < e; i++) {
                        const int64_t o = ob[i].observed >= (uint64_t)kRL ? kRL - 1 : (int64_t)ob[i].observed;
                        const int row = ob[i].strand ? b : 3 - b;
                        lp += tab[(row * kRL + u) * kRL + o] * ob[i].weight;
                    }
                    lp /= 10000000.0;
                    if (u == lo || lp > best) {
                        best = lp;
                        c = u;
                    }
                }
            }
            return c == 0 ? 1 : (uint64_t)c;
        };
        if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
        threads = (int)std::max<int64_t>(1, std::min<int64_t>(threads, (n + 1023) / 1024));
        std::atomic<int64_t> next{1};
        auto work = [&]() {
            for (int64_t k0; (k0 = next.fetch_add(1024)) < n;)
                for (int64_t k = k0; k < std::min<int64_t>(n, k0 + 1024); k++) repeatCounts[k - 1] = node_count(k);
        };
        mp::parallel_run(threads, [&](int) { work(); });
    } catch (const std::bad_alloc &) {
        mp::set_error("mp_poa_estimate_repeat_counts: out of host memory");
        return MP_ENOMEM;
    }
    return MP_OK;
}

extern "C" mp_status mp_poa_anchor_alignments(const mp_poa *poa, const int64_t *poaToConsensusMap, int64_t nReads,
                                              const double *anchorParams, int64_t nAnchorParams, int64_t *counts,
                                              int64_t *anchors, int64_t capacity) {
    if (!poa || nReads < 0 || !counts || !anchorParams || nAnchorParams < 2 || (nAnchorParams & 1)) {
        mp::set_error("mp_poa_anchor_alignments: bad argument (anchor parameters come in threshold/expansion pairs)");
        return MP_EINVAL;
    }
    const std::vector<mp_aligned_pair> &op = poa->obsPairs;
    const size_t nseg = poa->obsSeg.size();
    for (size_t sg = 0; sg < nseg; sg++)
        if (poa->obsSeg[sg].first < 0 || poa->obsSeg[sg].first >= nReads) {
            mp::set_error("mp_poa_anchor_alignments: an observation's read number is not below nReads");
            return MP_EINVAL;
        }
    try {
        // every read's observations in the reference's visiting order: node
        // ascending, then the node's observation order (append order)
        std::vector<int64_t> rStart(nReads + 1, 0);
        for (size_t sg = 0; sg < nseg; sg++) {
            const int64_t e = sg + 1 < nseg ? poa->obsSeg[sg + 1].second : (int64_t)op.size();
            rStart[poa->obsSeg[sg].first + 1] += e - poa->obsSeg[sg].second;
        }
        for (int64_t r = 0; r < nReads; r++) rStart[r + 1] += rStart[r];
        std::vector<int64_t> idx(op.size());  // tuple indices grouped by read, append order within a read
        {
            std::vector<int64_t> at(rStart.begin(), rStart.end() - 1);
            for (size_t sg = 0; sg < nseg; sg++) {
                const int64_t e = sg + 1 < nseg ? poa->obsSeg[sg + 1].second : (int64_t)op.size();
                int64_t &w = at[poa->obsSeg[sg].first];
                for (int64_t i = poa->obsSeg[sg].second; i < e; i++) idx[w++] = i;
            }
        }
        std::vector<std::vector<int64_t>> out((size_t)nReads);
        const int64_t nNodes = poa->nodes();
        const double thr0 = anchorParams[0];
        auto per_read = [&](int64_t r) {
            int64_t *a = idx.data() + rStart[r], *e = idx.data() + rStart[r + 1];
            std::stable_sort(a, e, [&](int64_t u, int64_t v) { return op[u].x < op[v].x; });
            std::vector<int64_t> &lst = out[r];
            for (int64_t *q = a; q < e; q++) {
                const mp_aligned_pair &t = op[*q];
                const int64_t node = (int64_t)t.x + 1;
                if (node < 1 || node >= nNodes) continue;
                const int64_t ci = poaToConsensusMap ? poaToConsensusMap[node - 1] : node - 1;
                if (ci == -1) continue;
                const double nw = (double)t.weight / 10000000.0;
                if (!(nw > thr0)) continue;
                int64_t expansion = (int64_t)anchorParams[1];
                for (int64_t k = 2; k < nAnchorParams; k += 2) {
                    if (nw >= anchorParams[k])
                        expansion = (int64_t)anchorParams[k + 1];
                    else
                        break;
                }
                const size_t L = lst.size();
                if (L == 0 || (lst[L - 3] < ci && lst[L - 2] < (int64_t)t.y)) {
                    lst.push_back(ci);
                    lst.push_back(t.y);
                    lst.push_back(expansion);
                }
            }
        };
        const int nth = (int)std::max<int64_t>(1, std::min<int64_t>(std::thread::hardware_concurrency(), nReads));
        std::atomic<int64_t> next{0};
        auto work = [&]() {
            for (int64_t r; (r = next.fetch_add(1)) < nReads;) per_read(r);
        };
        mp::parallel_run(nth, [&](int) { work(); });
        int64_t total = 0;
        for (int64_t r = 0; r < nReads; r++) {
            counts[r] = (int64_t)out[r].size() / 3;
            total += counts[r];
        }
        if (anchors) {
            if (capacity < total) {
                mp::set_error("mp_poa_anchor_alignments: capacity below the " + std::to_string(total) + " anchors");
                return MP_EINVAL;
            }
            int64_t w = 0;
            for (int64_t r = 0; r < nReads; r++) {
                std::copy(out[r].begin(), out[r].end(), anchors + w);
                w += (int64_t)out[r].size();
            }
        }
    } catch (const std::bad_alloc &) {
        mp::set_error("mp_poa_anchor_alignments: out of host memory");
        return MP_ENOMEM;
    }
    return MP_OK;
}

namespace {

// The reference's logAdd (pairwiseAligner.c:282-299): log(e^x + e^y) through a
// 4-piece cubic of log(1 + e^d), coefficients written as float literals (so
// rounded to float, then used in double arithmetic), cut at d >= 7.5.
inline double log1pexp_cubic(double d) {
    if (d <= 1.0f) return ((-0.009350833524763f * d + 0.130659527668286f) * d + 0.498799810682272f) * d + 0.693203116424741f;
    if (d <= 2.5f) return ((-0.014532321752540f * d + 0.139942324101744f) * d + 0.495635523139337f) * d + 0.692140569840976f;
    if (d <= 4.5f) return ((-0.004605031767994f * d + 0.063427417320019f) * d + 0.695956496475118f) * d + 0.514272634594009f;
    return ((-0.000458661602210f * d + 0.009695946122598f) * d + 0.930734667215156f) * d + 0.168037164329057f;
}

inline double log_add(double x, double y) {
    const double kNeg = -std::numeric_limits<double>::infinity();
    if (x < y) return (x == kNeg || y - x >= 7.5) ? y : log1pexp_cubic(y - x) + x;
    return (y == kNeg || x - y >= 7.5) ? x : log1pexp_cubic(x - y) + y;
}

// getMaxWeight (polisher.c:1025-1038): the heaviest non-reference entry
// (last of equals), unless the reference entry scaled by the penalty beats it.
int64_t max_weight(const double *w, int64_t n, int64_t ref, double penalty) {
    double mw = 0.0;
    int64_t mi = -1;
    for (int64_t j = 0; j < n; j++)
        if (j != ref && w[j] >= mw) {
            mw = w[j];
            mi = j;
        }
    const double wr = ref >= 0 && ref < n ? w[ref] : 0.0;
    return (wr * penalty > mw || mi < 0) ? ref : mi;
}

}  // namespace

extern "C" mp_status mp_poa_consensus(const mp_poa *poa, int32_t useRunLengthEncoding, double referenceBasePenalty,
                                      char *out, int64_t outCap, int64_t *outLen, int64_t *poaToConsensusMap) {
    if (!poa || !outLen) {
        mp::set_error("mp_poa_consensus: null argument");
        return MP_EINVAL;
    }
    const bool rle = useRunLengthEncoding != 0;
    const double kNeg = -std::numeric_limits<double>::infinity();
    try {
        const int64_t n = poa->nodes();
        if (poa->consGen == poa->gen.load() && poa->consRle == (rle ? 1 : 0) &&
            poa->consPenalty == referenceBasePenalty) {  // computed for this graph and these arguments
            const int64_t total = (int64_t)poa->consString.size();
            *outLen = total;
            if (poaToConsensusMap && n > 1)
                std::memcpy(poaToConsensusMap, poa->consMap.data(), sizeof(int64_t) * (size_t)(n - 1));
            if (!out || outCap < total) {
                mp::set_error("mp_poa_consensus: output buffer too small (need " + std::to_string(total) + ")");
                return MP_EINVAL;
            }
            if (total) std::memcpy(out, poa->consString.data(), (size_t)total);
            return MP_OK;
        }
        const int64_t mrc = poa->maxRC;
        auto ins_weight = [&](int32_t q) { return poa->ins[q].wF + poa->ins[q].wR; };
        auto del_weight = [&](int32_t q) { return poa->del[q].wF + poa->del[q].wR; };
        std::vector<double> totalOut(n, 0.0), fwd(n + 1, kNeg), matchFwd(n, 0.0);
        fwd[0] = 0.0;
        // deletions arriving at each node index (from node i, length L -> i + L + 1), in node then list order
        std::vector<std::vector<int32_t>> incoming(n + 1);
        for (int64_t i = 0; i < n; i++)
            for (int32_t q = poa->delHead[i]; q >= 0; q = poa->del[q].next)
                if (i + poa->del[q].len + 1 <= n) incoming[i + poa->del[q].len + 1].push_back(q);
        std::vector<int64_t> delFrom(poa->del.size(), 0);
        for (int64_t i = 0; i < n; i++)
            for (int32_t q = poa->delHead[i]; q >= 0; q = poa->del[q].next) delFrom[q] = i;
        // forward pass (polisher.c:1073-1136)
        for (int64_t i = 0; i < n; i++) {
            double indel = 0.0;
            for (int32_t q = poa->insHead[i]; q >= 0; q = poa->ins[q].next) indel += ins_weight(q);
            for (int32_t q = poa->delHead[i]; q >= 0; q = poa->del[q].next) indel += del_weight(q);
            double mt = 0.0;
            if (i == 0) {
                if (n == 1) {
                    mt = 1.0;
                } else {
                    for (int64_t j = 1; j < n; j++)
                        for (int k = 0; k < kAlphabet; k++) mt += poa->baseW[j * kAlphabet + k];
                    mt /= (double)(n - 1);
                    mt -= indel;
                }
            } else {
                for (int k = 0; k < kAlphabet; k++) mt += poa->baseW[i * kAlphabet + k];
                mt -= indel;
            }
            mt = mt <= 0.0 ? 0.0001 : mt;
            totalOut[i] = mt + indel;
            for (int32_t q = poa->insHead[i]; q >= 0; q = poa->ins[q].next)
                fwd[i + 1] = log_add(fwd[i + 1], fwd[i] + std::log(ins_weight(q) / totalOut[i]));
            for (int32_t q = poa->delHead[i]; q >= 0; q = poa->del[q].next) {
                const int64_t t = i + poa->del[q].len + 1;
                if (t <= n) fwd[t] = log_add(fwd[t], fwd[i] + std::log(del_weight(q) / totalOut[i]));
            }
            matchFwd[i] = fwd[i] + std::log(mt / totalOut[i]);
            fwd[i + 1] = log_add(fwd[i + 1], matchFwd[i]);
        }
        // greedy traceback from the end (polisher.c:1138-1245)
        std::vector<int64_t> map((size_t)std::max<int64_t>(n - 1, 0), -1);
        std::vector<std::pair<const char *, int64_t>> pieces;  // (base run or insert) in reverse order
        std::string store;                                    // backing for expanded runs, filled later
        std::vector<std::pair<char, int64_t>> runs;           // one entry per piece: base and count, or insert
        std::vector<int32_t> pieceIns;                        // insert index of a piece, -1 for a node base
        int64_t running = 0;
        char prev = '-';
        for (int64_t i = n; i > 0;) {
            if (i < n) {
                const int bsym = base_symbol(poa->nodeBase[i]);
                const int64_t mb = max_weight(&poa->baseW[i * kAlphabet], kAlphabet, bsym, referenceBasePenalty);
                const char base = "ACGTN"[mb];
                int64_t rc = 1;
                if (rle) {
                    rc = max_weight(&poa->rcW[i * mrc], mrc, (int64_t)std::min<uint64_t>(poa->nodeRC[i], INT64_MAX),
                                    referenceBasePenalty);
                    rc = rc == 0 ? 1 : rc;
                    if (prev != base) map[i - 1] = running++;
                    prev = base;
                } else {
                    map[i - 1] = running++;
                }
                runs.emplace_back(base, rc);
                pieceIns.push_back(-1);
            }
            double maxIns = kNeg, totIns = kNeg;
            int32_t bestIns = -1;
            for (int32_t q = poa->insHead[i - 1]; q >= 0; q = poa->ins[q].next) {
                const double pr = std::log(ins_weight(q) / totalOut[i - 1]) + fwd[i - 1];
                if (pr > maxIns) {
                    maxIns = pr;
                    bestIns = q;
                }
                totIns = log_add(totIns, pr);
            }
            double maxDel = kNeg, totDel = kNeg;
            int32_t bestDel = -1;
            for (int32_t q : incoming[i]) {
                const int64_t src = delFrom[q];
                const double pr = std::log(del_weight(q) / totalOut[src]) + fwd[src];
                if (pr > maxDel) {
                    maxDel = pr;
                    bestDel = q;
                }
                totDel = log_add(totDel, pr);
            }
            if (matchFwd[i - 1] >= totDel && matchFwd[i - 1] >= totIns) {
                i--;
            } else if (totIns >= totDel) {
                const InsEntry &E = poa->ins[bestIns];
                runs.emplace_back('\0', 0);
                pieceIns.push_back(bestIns);
                if (rle) {
                    const char last = poa->insB[E.sym + E.len - 1];
                    running += E.len + (last != prev ? 0 : -1);
                    prev = poa->insB[E.sym];
                } else {
                    running += E.len;
                }
                i--;
            } else {
                i -= poa->del[bestDel].len + 1;
            }
        }
        // the pieces were collected from the end: expand them front to back
        int64_t total = 0;
        for (size_t q = 0; q < runs.size(); q++) {
            if (pieceIns[q] < 0) {
                total += runs[q].second;
            } else {
                const InsEntry &E = poa->ins[pieceIns[q]];
                for (int64_t t = 0; t < E.len; t++) total += (int64_t)poa->insC[E.sym + t];
            }
        }
        *outLen = total;
        std::string expanded;
        expanded.reserve((size_t)total);
        for (size_t q = runs.size(); q-- > 0;) {
            if (pieceIns[q] < 0) {
                expanded.append((size_t)runs[q].second, runs[q].first);
            } else {
                const InsEntry &E = poa->ins[pieceIns[q]];
                for (int64_t t = 0; t < E.len; t++)
                    expanded.append((size_t)poa->insC[E.sym + t], poa->insB[E.sym + t]);
            }
        }
        // consensus length in run-length (or plain) symbols, then the map from the string's start
        int64_t clen = (int64_t)expanded.size();
        if (rle) {
            clen = 0;
            for (size_t t = 0; t < expanded.size(); t++)
                if (t == 0 || expanded[t] != expanded[t - 1]) clen++;
        }
        for (int64_t i = 0; i + 1 < n; i++) map[i] = map[i] == -1 ? -1 : clen - 1 - map[i];
        if (poaToConsensusMap && n > 1) std::memcpy(poaToConsensusMap, map.data(), sizeof(int64_t) * (size_t)(n - 1));
        poa->consString.swap(expanded);
        poa->consMap.swap(map);
        poa->consGen = poa->gen.load();
        poa->consRle = rle ? 1 : 0;
        poa->consPenalty = referenceBasePenalty;
        if (!out || outCap < total) {
            mp::set_error("mp_poa_consensus: output buffer too small (need " + std::to_string(total) + ")");
            return MP_EINVAL;
        }
        if (total) std::memcpy(out, poa->consString.data(), (size_t)total);
    } catch (const std::bad_alloc &) {
        mp::set_error("mp_poa_consensus: out of host memory");
        return MP_ENOMEM;
    }
    return MP_OK;
}

extern "C" mp_status mp_poa_scores(const mp_poa *poa, double *scores) {
    if (!poa || !scores) {
        mp::set_error("mp_poa_scores: null argument");
        return MP_EINVAL;
    }
    double match = 0.0, other = 0.0, ins = 0.0, del = 0.0;
    for (int64_t i = 0; i < poa->nodes(); i++) {
        const int rs = base_symbol(poa->nodeBase[i]);
        match += poa->baseW[i * kAlphabet + rs];
        for (int j = 0; j < kAlphabet; j++)
            if (j != rs) other += poa->baseW[i * kAlphabet + j];
        for (int32_t q = poa->insHead[i]; q >= 0; q = poa->ins[q].next)
            ins += (poa->ins[q].wF + poa->ins[q].wR) * (double)poa->ins[q].len;
        for (int32_t q = poa->delHead[i]; q >= 0; q = poa->del[q].next)
            del += (poa->del[q].wF + poa->del[q].wR) * (double)poa->del[q].len;
    }
    scores[0] = match;
    scores[1] = other;
    scores[2] = ins;
    scores[3] = del;
    return MP_OK;
}

extern "C" mp_status mp_poas_estimate_repeat_counts(const mp_poa *const *poas, int64_t nPoas, const mp_model *m,
                                                    const mp_rle_string *reads, const int32_t *readStrand,
                                                    const int64_t *readStart, uint64_t *const *repeatCounts,
                                                    int32_t threads) {
    if (!poas || nPoas < 0 || !m || (nPoas > 0 && (!readStart || !repeatCounts))) {
        mp::set_error("mp_poas_estimate_repeat_counts: bad argument");
        return MP_EINVAL;
    }
    for (int64_t g = 0; g < nPoas; g++)
        if (!poas[g] || !repeatCounts[g] || readStart[g + 1] < readStart[g]) {
            mp::set_error("mp_poas_estimate_repeat_counts: bad graph " + std::to_string(g));
            return MP_EINVAL;
        }
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    threads = (int)std::max<int64_t>(1, std::min<int64_t>(threads, nPoas));
    std::atomic<int64_t> next{0};
    std::atomic<int64_t> badGraph{-1};
    std::vector<std::string> errs((size_t)std::max<int64_t>(nPoas, 1));
    std::vector<mp_status> sts((size_t)std::max<int64_t>(nPoas, 1), MP_OK);
    auto work = [&]() {
        for (int64_t g; (g = next.fetch_add(1)) < nPoas;) {
            const int64_t r0 = readStart[g];
            sts[g] = mp_poa_estimate_repeat_counts(poas[g], m, reads + r0, readStrand + r0, readStart[g + 1] - r0,
                                                   repeatCounts[g], 1);
            if (sts[g] != MP_OK) {
                errs[g] = mp_last_error();
                int64_t none = -1;
                badGraph.compare_exchange_strong(none, g);
            }
        }
    };
    try {
        mp::parallel_run(threads, [&](int) { work(); });
    } catch (const std::bad_alloc &) {
        mp::set_error("mp_poas_estimate_repeat_counts: out of host memory");
        return MP_ENOMEM;
    }
    if (badGraph.load() >= 0) {
        mp::set_error("graph " + std::to_string(badGraph.load()) + ": " + errs[badGraph.load()]);
        return sts[badGraph.load()];
    }
    return MP_OK;
}

