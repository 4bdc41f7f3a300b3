// exact_dp.cpp -- independent exact / modular CPU count of corner-to-corner
// simple paths on an (n+1)x(n+1) vertex grid (OEIS A007764).
//
// TEST INFRASTRUCTURE ONLY (the second, independent oracle).  It restates
// the semantics of the reference's vertex-degree / component frontier DP,
// src/cpp/oneesan_frontier.cpp:87-245: edges taken in row-major order
// (right edge, then down edge, per vertex), per frontier vertex a degree and
// a component label, per component a start/target flag pair, and a global
// "done" flag.  The reference uses boost cpp_int; this restatement uses
// unsigned __int128 (exact for n <= 12) or uint32 residues mod p.
//
// It shares no code and no state encoding with the Grid-FP oracle
// (gridfp_oracle.c), so agreement between the two pins the Grid-FP semantics
// end to end.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

struct St {
    bool done = false;
    std::vector<uint8_t> deg, comp;  // per frontier slot
    std::vector<uint8_t> flags;      // per component id (index 0 unused)
};

// Relabel components by first appearance (oneesan_frontier.cpp:35-51).
void canon(St& s) {
    uint8_t remap[64] = {0};
    uint8_t nf[64] = {0};
    uint8_t next = 1;
    for (auto& c : s.comp) {
        if (!c) continue;
        if (!remap[c]) { remap[c] = next; nf[next] = s.flags[c]; ++next; }
        c = remap[c];
    }
    s.flags.assign(next, 0);
    for (uint8_t c = 1; c < next; ++c) s.flags[c] = nf[c];
}

std::string key_of(const St& s) {
    std::string k;
    k.reserve(2 + 2 * s.deg.size() + s.flags.size());
    k.push_back(char(s.done));
    k.push_back(char(s.deg.size()));
    for (size_t i = 0; i < s.deg.size(); ++i) { k.push_back(char(s.deg[i])); k.push_back(char(s.comp[i])); }
    for (size_t i = 1; i < s.flags.size(); ++i) k.push_back(char(s.flags[i]));
    return k;
}

St decode(const std::string& k) {
    St s;
    size_t p = 0;
    s.done = k[p++] != 0;
    size_t n = (uint8_t)k[p++];
    s.deg.resize(n); s.comp.resize(n);
    uint8_t mx = 0;
    for (size_t i = 0; i < n; ++i) { s.deg[i] = (uint8_t)k[p++]; s.comp[i] = (uint8_t)k[p++]; mx = std::max(mx, s.comp[i]); }
    s.flags.assign(mx + 1, 0);
    for (uint8_t c = 1; c <= mx; ++c) s.flags[c] = (uint8_t)k[p++];
    return s;
}

template <class C, class Add>
C count_paths(int n, Add add) {
    const int w = n + 1, V = w * w, start = 0, target = V - 1;
    std::vector<std::pair<int, int>> edges;
    for (int r = 0; r < w; ++r)
        for (int c = 0; c < w; ++c) {
            int u = r * w + c;
            if (c + 1 < w) edges.push_back({u, u + 1});
            if (r + 1 < w) edges.push_back({u, u + w});
        }
    const int E = (int)edges.size();
    std::vector<int> first(V, 1 << 30), last(V, -1);
    for (int i = 0; i < E; ++i) {
        auto [u, v] = edges[i];
        first[u] = std::min(first[u], i); first[v] = std::min(first[v], i);
        last[u] = std::max(last[u], i); last[v] = std::max(last[v], i);
    }
    auto active = [&](int e) {
        std::vector<int> a;
        for (int v = 0; v < V; ++v) if (first[v] < e && last[v] >= e) a.push_back(v);
        return a;
    };
    std::unordered_map<std::string, C> cur, nxt;
    St init;
    init.flags.assign(1, 0);
    cur[key_of(init)] = C(1);
    std::vector<int> before = active(0);
    for (int ei = 0; ei < E; ++ei) {
        std::vector<int> after = active(ei + 1);
        auto [u, v] = edges[ei];
        std::vector<int> work = before;
        for (int x : {u, v}) if (std::find(work.begin(), work.end(), x) == work.end()) work.push_back(x);
        std::vector<int> wpos(V, -1), apos(V, -1);
        for (int i = 0; i < (int)work.size(); ++i) wpos[work[i]] = i;
        for (int i = 0; i < (int)after.size(); ++i) apos[after[i]] = i;
        nxt.clear();
        for (auto const& kv : cur) {
            St old = decode(kv.first);
            St base;
            base.done = old.done;
            base.deg.assign(work.size(), 0);
            base.comp.assign(work.size(), 0);
            base.flags = old.flags;
            for (size_t i = 0; i < before.size(); ++i) {
                base.deg[wpos[before[i]]] = old.deg[i];
                base.comp[wpos[before[i]]] = old.comp[i];
            }
            for (int take = 0; take < 2; ++take) {
                St s = base;
                const int iu = wpos[u], iv = wpos[v];
                if (take) {
                    if (s.done) continue;
                    const uint8_t mu = (u == start || u == target) ? 1 : 2;
                    const uint8_t mv = (v == start || v == target) ? 1 : 2;
                    if (s.deg[iu] >= mu || s.deg[iv] >= mv) continue;
                    ++s.deg[iu]; ++s.deg[iv];
                    uint8_t a = s.comp[iu], b = s.comp[iv];
                    uint8_t term = (uint8_t)(((u == start || v == start) ? 1 : 0) | ((u == target || v == target) ? 2 : 0));
                    if (!a && !b) {
                        uint8_t q = (uint8_t)s.flags.size();
                        s.flags.push_back(term);
                        s.comp[iu] = s.comp[iv] = q;
                    } else if (!a || !b) {
                        uint8_t q = a ? a : b;
                        s.comp[iu] = s.comp[iv] = q;
                        s.flags[q] |= term;
                    } else {
                        if (a == b) continue;  // would close a cycle
                        uint8_t keep = std::min(a, b), kill = std::max(a, b);
                        s.flags[keep] |= s.flags[kill];
                        for (auto& c : s.comp) if (c == kill) c = keep;
                    }
                }
                bool ok = true;
                for (int x : work) {
                    if (apos[x] != -1) continue;
                    int ix = wpos[x];
                    bool terminal = (x == start || x == target);
                    if ((terminal && s.deg[ix] != 1) || (!terminal && s.deg[ix] != 0 && s.deg[ix] != 2)) { ok = false; break; }
                    uint8_t q = s.comp[ix];
                    s.comp[ix] = 0;
                    if (q) {
                        bool alive = false;
                        for (uint8_t c : s.comp) if (c == q) { alive = true; break; }
                        if (!alive) {
                            if (s.flags[q] == 3) s.done = true;
                            else { ok = false; break; }
                        }
                    }
                }
                if (!ok) continue;
                St out;
                out.done = s.done;
                out.deg.resize(after.size());
                out.comp.resize(after.size());
                out.flags = s.flags;
                for (size_t i = 0; i < after.size(); ++i) {
                    out.deg[i] = s.deg[wpos[after[i]]];
                    out.comp[i] = s.comp[wpos[after[i]]];
                }
                canon(out);
                auto it = nxt.find(key_of(out));
                if (it == nxt.end()) nxt.emplace(key_of(out), kv.second);
                else it->second = add(it->second, kv.second);
            }
        }
        cur.swap(nxt);
        before.swap(after);
    }
    C ans(0);
    for (auto const& kv : cur) if (decode(kv.first).done) ans = add(ans, kv.second);
    return ans;
}

}  // namespace

extern "C" {

// Exact count as a decimal string into buf (n <= 12 fits in 128 bits).
// Returns 0 on success, nonzero on bad arguments.
int xdp_count_exact(int n, char* buf, int cap) {
    if (n < 1 || n > 12 || cap < 48) return 1;
    typedef unsigned __int128 u128;
    u128 v = count_paths<u128>(n, [](u128 a, u128 b) { return a + b; });
    char tmp[64];
    int len = 0;
    if (!v) tmp[len++] = '0';
    while (v) { tmp[len++] = char('0' + int(v % 10)); v /= 10; }
    for (int i = 0; i < len; ++i) buf[i] = tmp[len - 1 - i];
    buf[len] = 0;
    return 0;
}

// Count mod p (2 <= p < 2^32).
uint32_t xdp_count_mod(int n, uint32_t p, int* err) {
    *err = 0;
    if (n < 1 || n > 14 || p < 2) { *err = 1; return 0; }
    uint32_t one = 1 % p;
    uint32_t r = count_paths<uint32_t>(n, [p](uint32_t a, uint32_t b) {
        uint64_t s = uint64_t(a) + b;
        return uint32_t(s >= p ? s - p : s);
    });
    (void)one;
    return r;
}

}  // extern "C"
