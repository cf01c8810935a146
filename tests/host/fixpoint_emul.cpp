// Host emulation of the forward solver's fixed-point iteration (no tiles): the device's local heap
// replay (eik_stencil.h, replay()) is applied to every node of a grid, in random orders, until no
// node changes, and the converged traveltimes are compared bit for bit with the oracle's heap FMM
// (og_fmm_record) on the same model and source.  Orders:
//   gs    -- Gauss-Seidel: a random permutation per pass, each result written at once;
//   jac   -- Jacobi: every node of a pass reads the previous pass's state;
//   torn  -- Jacobi where each neighbour read takes the value from the previous pass and its pop key
//            (K, S, D) from the pass before, half of the time (a reader that sees a half-written
//            (value, key) tuple of a node being written concurrently);
//   part  -- Gauss-Seidel over a random half of the nodes per pass (partial, band-like schedules).
// A unique, reachable fixed point must give the heap's answer under every order.  The kernel's own
// wake-up rule and storage are emulated as well (run_filtered, below).  Models: 3D
// spherical with a radial velocity (many near-ties, like the 120x150x150 benchmark grid) plus a
// smooth perturbation, and Cartesian 2D/3D smooth models; several off-node sources each.
#include "eik_stencil.h"
extern "C" {
#include "eik_oracle.h"
}
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>
using namespace eik;

struct State {
    std::vector<double> tt, K, S;
    std::vector<uint8_t> D;
};

static bool same(double a, double b) { return memcmp(&a, &b, sizeof(double)) == 0; }

struct Model {
    og_grid g;
    std::vector<double> vel;
    int nd;
    int64_t N;
};

static Model make_model(int kind, int n0, int n1, int n2, unsigned seed) {
    Model M{};
    M.g.kind = kind;
    M.nd = (kind == OG_CART2D || kind == OG_SPH2D) ? 2 : 3;
    M.g.ndim = M.nd;
    M.g.n[0] = n0;
    M.g.n[1] = n1;
    M.g.n[2] = M.nd == 3 ? n2 : 1;
    M.N = M.g.n[0] * M.g.n[1] * M.g.n[2];
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    double ph[6][4];
    for (auto& m : ph)
        for (double& x : m) x = U(rng);
    if (kind == OG_SPH3D) {
        M.g.d[0] = 2.0;
        M.g.d[1] = 0.1;
        M.g.d[2] = 0.1;
        M.g.o[0] = 6371.0 - 2.0 * (n0 - 1);
        M.g.o[1] = 82.55;
        M.g.o[2] = 10.0;
    } else {
        for (int d = 0; d < 3; d++) {
            M.g.d[d] = 1.0;
            M.g.o[d] = 0.0;
        }
    }
    M.vel.resize(M.N);
    for (int64_t l = 0; l < M.N; l++) {
        const int64_t i = l % n0, j = (l / n0) % n1, k = M.nd == 3 ? l / ((int64_t)n0 * n1) : 0;
        double pert = 0.0;
        for (auto& m : ph)
            pert += sin(2 * M_PI * (m[0] * 3 * i / n0 + m[1] * 3 * j / n1 + m[2] * 3 * k / (n2 > 1 ? n2 : 1) + m[3]));
        pert /= 6.0;
        if (kind == OG_SPH3D) {
            const double r = M.g.o[0] + M.g.d[0] * i;
            M.vel[l] = (5.8 + 2.2 * (6371.0 - r) / 238.0) * (1.0 + 0.03 * pert);
        } else {
            M.vel[l] = 3.0 + 1.2 * pert + 0.02 * j;
        }
    }
    return M;
}

static void node_ijk(const Model& M, int64_t l, int* g) {
    g[0] = (int)(l % M.g.n[0]);
    g[1] = (int)((l / M.g.n[0]) % M.g.n[1]);
    g[2] = (int)(l / (M.g.n[0] * M.g.n[1]));
}
static int64_t lin3(const Model& M, const int* g) { return g[0] + M.g.n[0] * (g[1] + M.g.n[1] * (int64_t)g[2]); }

// one replay of node l reading `val` for values and `key` for pop keys
static void update(const Model& M, const State& val, const State& key, int64_t l, double* nv, Key<double>* nk) {
    const int nd = M.nd;
    int g[3];
    node_ijk(M, l, g);
    double v1[6], v2[6];
    Key<double> k1[6], k2[6];
    unsigned f1 = 0, f2 = 0, ok1 = 0, ok2 = 0;
    for (int k = 0; k < 2 * nd; k++) {
        const int ax = k >> 1, sh = (k & 1) ? -1 : 1;
        int a[3] = {g[0], g[1], g[2]}, b[3] = {g[0], g[1], g[2]};
        a[ax] += sh;
        b[ax] += 2 * sh;
        const bool in1 = a[ax] >= 0 && a[ax] <= M.g.n[ax] - 1;
        bool in2 = in1;
        for (int d = 0; d < nd; d++) in2 = in2 && a[d] >= 1 && a[d] <= M.g.n[d] - 2;
        double ra = INFINITY, rb = INFINITY;
        Key<double> ka{INFINITY, -INFINITY, 0}, kb{INFINITY, -INFINITY, 0};
        if (in1) {
            const int64_t q = lin3(M, a);
            ra = val.tt[q];
            ka = Key<double>{key.K[q], key.S[q], (int)key.D[q]};
        }
        if (in2) {
            const int64_t q = lin3(M, b);
            rb = val.tt[q];
            kb = Key<double>{key.K[q], key.S[q], (int)key.D[q]};
        }
        ok1 |= (unsigned)in1 << k;
        ok2 |= (unsigned)in2 << k;
        f1 |= (unsigned)std::signbit(ra) << k;
        f2 |= (unsigned)std::signbit(rb) << k;
        v1[k] = fabs(ra);
        v2[k] = fabs(rb);
        k1[k] = ka;
        k2[k] = kb;
    }
    Metric<double> m;
    if (M.g.kind == OG_SPH3D) {
        const double r = fma((double)g[0], M.g.d[0], M.g.o[0]);
        m.dh[0] = M.g.d[0];
        m.dh[1] = r * (M.g.d[1] * (M_PI / 180.0));
        m.dh[2] = r * sin(fma((double)g[1], M.g.d[1], M.g.o[1]) * (M_PI / 180.0)) * (M.g.d[2] * (M_PI / 180.0));
    } else {
        m.dh[0] = m.dh[1] = m.dh[2] = M.g.d[0];
    }
    metric_coefs(m, nd);
    int code[3];
    const double s = 1.0 / M.vel[l];
    if (nd == 2)
        *nv = replay<2, double, false>(v1, v2, k1, k2, f1, f2, ok1, ok2, s, m, code, nk);
    else
        *nv = replay<3, double, false>(v1, v2, k1, k2, f1, f2, ok1, ok2, s, m, code, nk);
}

static bool write(State& st, int64_t l, double nv, const Key<double>& nk) {
    const bool vchg = nv == nv && !same(nv, st.tt[l]) && !(nv >= INFINITY && st.tt[l] >= INFINITY);
    const bool kchg = nv < INFINITY && (!same(nk.K, st.K[l]) || !same(nk.S, st.S[l]) || nk.D != st.D[l]);
    if (!(vchg || kchg)) return false;
    st.tt[l] = nv;
    st.K[l] = nk.D ? nk.K : nv;
    st.S[l] = nk.D ? nk.S : -INFINITY;
    st.D[l] = (uint8_t)nk.D;
    return true;
}

// The kernel's storage and wake-up rule, at node granularity.  Memory holds tt, D and the cascade key
// pools K, S; a node is re-evaluated only when a stencil neighbour changed with a key (old or new)
// not above its own, and a reader racing a writer may see each of the four fields from before or
// after the write.  keep_keys = 0: the pools are written only for cascade nodes (D != 0) and hold
// stale values (here: another solve's keys) elsewhere; 1: they start as (+inf, -inf), an unreached
// node's key, and every write stores K = |t|, S = -inf for ordinary nodes, so every field a reader
// can see belongs to the node's old or new state.
struct Mem {
    std::vector<double> tt, K, S;
    std::vector<uint8_t> D;
};
static long run_filtered(const Model& M, const std::vector<double>& heap, const std::vector<int64_t>& init,
                         bool keep_keys, bool verify, unsigned seed, long* passes, double* maxrel) {
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    const int nd = M.nd;
    Mem cur;
    cur.tt.assign(M.N, INFINITY);
    cur.D.assign(M.N, 0);
    cur.K.resize(M.N);
    cur.S.resize(M.N);
    for (int64_t l = 0; l < M.N; l++) {
        // keep_keys: the pools start as an unreached node's key; otherwise they hold leftovers of
        // another solve
        cur.K[l] = keep_keys ? INFINITY : 0.5 * heap[l] * U(rng);
        cur.S[l] = -INFINITY;
    }
    std::vector<char> isinit(M.N, 0);
    std::vector<double> trig(M.N, INFINITY);
    for (int64_t l : init) {
        cur.tt[l] = -heap[l];
        isinit[l] = 1;
    }
    auto key_of = [&](const Mem& m, int64_t q) {
        return m.D[q] ? Key<double>{m.K[q], m.S[q], (int)m.D[q]} : Key<double>{fabs(m.tt[q]), -INFINITY, 0};
    };
    // neighbours of the init nodes start pending
    auto mark = [&](int64_t l, double t) {
        int g[3];
        node_ijk(M, l, g);
        for (int ax = 0; ax < nd; ax++)
            for (int dd = -2; dd <= 2; dd++) {
                if (!dd) continue;
                int q[3] = {g[0], g[1], g[2]};
                q[ax] += dd;
                if (q[ax] < 0 || q[ax] > M.g.n[ax] - 1) continue;
                const int64_t lq = lin3(M, q);
                if (isinit[lq]) continue;
                if (t < trig[lq]) trig[lq] = t;
            }
    };
    for (int64_t l : init) mark(l, 0.0);
    long p = 0;
    const long cap = 40000;
    std::vector<int64_t> work;
    bool verifying = false;
    for (; p < cap; p++) {
        work.clear();
        for (int64_t l = 0; l < M.N; l++)
            if (trig[l] < INFINITY) {
                // a trigger above the node's own key (as it is when the node is taken up) cannot
                // change its replay: dropped, as the kernel does
                if (trig[l] <= key_of(cur, l).K) work.push_back(l);
                trig[l] = INFINITY;
            }
        if (work.empty()) {
            // verify: one pass over every node; a pass that changes nothing proves a fixed point
            if (!verify || verifying) break;
            verifying = true;
            for (int64_t l = 0; l < M.N; l++)
                if (!isinit[l]) work.push_back(l);
        } else {
            verifying = false;
        }
        std::shuffle(work.begin(), work.end(), rng);
        const Mem pre = cur;  // memory at the start of the pass; writes of this pass race the reads
        for (int64_t l : work) {
            // gather with torn reads: each field of each neighbour from before or after this pass
            int g[3];
            node_ijk(M, l, g);
            double v1[6], v2[6];
            Key<double> k1[6], k2[6];
            unsigned f1 = 0, f2 = 0, ok1 = 0, ok2 = 0;
            auto rd = [&](int64_t q, double* v, Key<double>* kk) {
                const Mem& a = (rng() & 1) ? pre : cur;
                const Mem& b = (rng() & 1) ? pre : cur;
                const Mem& c = (rng() & 1) ? pre : cur;
                *v = a.tt[q];
                const int D = b.D[q];
                *kk = D ? Key<double>{c.K[q], c.S[q], D} : Key<double>{fabs(*v), -INFINITY, 0};
            };
            for (int k = 0; k < 2 * nd; k++) {
                const int ax = k >> 1, sh = (k & 1) ? -1 : 1;
                int a[3] = {g[0], g[1], g[2]}, b[3] = {g[0], g[1], g[2]};
                a[ax] += sh;
                b[ax] += 2 * sh;
                const bool in1 = a[ax] >= 0 && a[ax] <= M.g.n[ax] - 1;
                bool in2 = in1;
                for (int d = 0; d < nd; d++) in2 = in2 && a[d] >= 1 && a[d] <= M.g.n[d] - 2;
                double ra = INFINITY, rb = INFINITY;
                Key<double> ka{INFINITY, -INFINITY, 0}, kb{INFINITY, -INFINITY, 0};
                if (in1) rd(lin3(M, a), &ra, &ka);
                if (in2) rd(lin3(M, b), &rb, &kb);
                ok1 |= (unsigned)in1 << k;
                ok2 |= (unsigned)in2 << k;
                f1 |= (unsigned)std::signbit(ra) << k;
                f2 |= (unsigned)std::signbit(rb) << k;
                v1[k] = fabs(ra);
                v2[k] = fabs(rb);
                k1[k] = ka;
                k2[k] = kb;
            }
            Metric<double> m;
            if (M.g.kind == OG_SPH3D) {
                const double r = fma((double)g[0], M.g.d[0], M.g.o[0]);
                m.dh[0] = M.g.d[0];
                m.dh[1] = r * (M.g.d[1] * (M_PI / 180.0));
                m.dh[2] = r * sin(fma((double)g[1], M.g.d[1], M.g.o[1]) * (M_PI / 180.0)) * (M.g.d[2] * (M_PI / 180.0));
            } else {
                m.dh[0] = m.dh[1] = m.dh[2] = M.g.d[0];
            }
            metric_coefs(m, nd);
            int code[3];
            Key<double> nk;
            const double s = 1.0 / M.vel[l];
            const double nv = nd == 2 ? replay<2, double, false>(v1, v2, k1, k2, f1, f2, ok1, ok2, s, m, code, &nk)
                                      : replay<3, double, false>(v1, v2, k1, k2, f1, f2, ok1, ok2, s, m, code, &nk);
            const Key<double> ko = key_of(cur, l);
            const double t0 = cur.tt[l];
            const bool vchg = nv == nv && !same(nv, t0) && !(nv >= INFINITY && t0 >= INFINITY);
            const bool kchg = nv < INFINITY && (!same(nk.K, ko.K) || !same(nk.S, ko.S) || nk.D != ko.D);
            if (!(vchg || kchg)) continue;
            cur.tt[l] = nv;
            cur.D[l] = (uint8_t)nk.D;
            if (nk.D || keep_keys) {
                cur.K[l] = nk.D ? nk.K : nv;
                cur.S[l] = nk.D ? nk.S : -INFINITY;
            }
            mark(l, ko.K < nk.K ? ko.K : nk.K);
        }
    }
    *passes = p;
    if (p >= cap) return -1;
    long ndiff = 0;
    double mr = 0.0;
    for (int64_t l = 0; l < M.N; l++) {
        const double a = fabs(cur.tt[l]);
        if (!same(a, heap[l])) {
            ndiff++;
            mr = std::max(mr, fabs(a - heap[l]) / heap[l]);
        }
    }
    *maxrel = mr;
    if (ndiff && getenv("FIXPOINT_DIAG")) {
        // is the converged state a fixed point of the replay (then the wake-up rule missed nothing)?
        State st;
        st.tt = cur.tt;
        st.D = cur.D;
        st.K.resize(M.N);
        st.S.resize(M.N);
        for (int64_t l = 0; l < M.N; l++) {
            const Key<double> k = key_of(cur, l);
            st.K[l] = k.K;
            st.S[l] = k.S;
        }
        long notfixed = 0;
        for (int64_t l = 0; l < M.N; l++) {
            if (isinit[l]) continue;
            double nv;
            Key<double> nk;
            update(M, st, st, l, &nv, &nk);
            const Key<double> ko = key_of(cur, l);
            if (!same(nv, cur.tt[l]) || !same(nk.K, ko.K) || !same(nk.S, ko.S) || nk.D != ko.D) {
                if (notfixed < 3) {
                    int g[3];
                    node_ijk(M, l, g);
                    printf("  not a fixed point at (%d,%d,%d): stored %.17g key (%.17g, %.3g, %d), replay %.17g "
                           "key (%.17g, %.3g, %d), heap %.17g\n", g[0], g[1], g[2], cur.tt[l], ko.K, ko.S, ko.D, nv,
                           nk.K, nk.S, nk.D, heap[l]);
                }
                notfixed++;
            }
        }
        printf("  nodes not at a fixed point: %ld\n", notfixed);
    }
    return ndiff;
}

// returns nodes differing from the heap (bitwise), or -1 when the iteration did not converge
static long run(const Model& M, const std::vector<double>& heap, const std::vector<int64_t>& init, int mode,
                unsigned seed, long* passes, double* maxrel) {
    std::mt19937_64 rng(seed);
    State cur;
    cur.tt.assign(M.N, INFINITY);
    cur.K.assign(M.N, INFINITY);
    cur.S.assign(M.N, -INFINITY);
    cur.D.assign(M.N, 0);
    std::vector<char> isinit(M.N, 0);
    for (int64_t l : init) {
        cur.tt[l] = -heap[l];  // init nodes carry the sign bit
        isinit[l] = 1;
    }
    std::vector<int64_t> order;
    for (int64_t l = 0; l < M.N; l++)
        if (!isinit[l]) order.push_back(l);
    State prev = cur, prev2 = cur;
    const long cap = 40000;
    long p = 0;
    for (; p < cap; p++) {
        std::shuffle(order.begin(), order.end(), rng);
        bool changed = false;
        if (mode == 0 || mode == 3) {  // gs / part
            for (int64_t l : order) {
                if (mode == 3 && (rng() & 1)) continue;
                double nv;
                Key<double> nk;
                update(M, cur, cur, l, &nv, &nk);
                changed |= write(cur, l, nv, nk);
            }
            if (mode == 3 && !changed) {  // a quiet partial pass: confirm with a full one
                for (int64_t l : order) {
                    double nv;
                    Key<double> nk;
                    update(M, cur, cur, l, &nv, &nk);
                    changed |= write(cur, l, nv, nk);
                }
            }
        } else {  // jac / torn
            prev2 = prev;
            prev = cur;
            State keysrc = prev;
            if (mode == 2)
                for (int64_t l = 0; l < M.N; l++)
                    if (rng() & 1) {
                        keysrc.K[l] = prev2.K[l];
                        keysrc.S[l] = prev2.S[l];
                        keysrc.D[l] = prev2.D[l];
                    }
            for (int64_t l : order) {
                double nv;
                Key<double> nk;
                update(M, prev, keysrc, l, &nv, &nk);
                changed |= write(cur, l, nv, nk);
            }
            if (!changed && mode == 2) {  // torn reads may hide a change: confirm untorn
                for (int64_t l : order) {
                    double nv;
                    Key<double> nk;
                    update(M, cur, cur, l, &nv, &nk);
                    changed |= write(cur, l, nv, nk);
                }
            }
        }
        if (!changed) break;
    }
    *passes = p;
    if (p >= cap) return -1;
    long nd = 0;
    double mr = 0.0;
    for (int64_t l = 0; l < M.N; l++) {
        const double a = fabs(cur.tt[l]);
        if (!same(a, heap[l])) {
            nd++;
            mr = std::max(mr, fabs(a - heap[l]) / heap[l]);
        }
    }
    *maxrel = mr;
    return nd;
}

int main(int argc, char** argv) {
    const int nsrc = argc > 1 ? atoi(argv[1]) : 3;
    const int nseeds = argc > 2 ? atoi(argv[2]) : 2;
    const bool do_orders = argc <= 3 || strcmp(argv[3], "filtered") != 0;
    const bool do_filtered = argc <= 3 || strcmp(argv[3], "orders") != 0;
    const char* names[4] = {"gs", "jac", "torn", "part"};
    struct Case {
        const char* name;
        int kind, n0, n1, n2;
    } cases[] = {{"sph3d-radial", OG_SPH3D, 30, 36, 36}, {"cart3d", OG_CART3D, 28, 26, 22}, {"cart2d", OG_CART2D, 90, 70, 1}};
    long bad = 0, total = 0, unsafe_off = 0;
    for (const Case& c : cases) {
        const Model M = make_model(c.kind, c.n0, c.n1, c.n2, 7);
        std::mt19937_64 rng(99);
        std::uniform_real_distribution<double> U(0.1, 0.9);
        for (int s = 0; s < nsrc; s++) {
            double src[3];
            for (int d = 0; d < M.nd; d++) src[d] = M.g.o[d] + U(rng) * (M.g.n[d] - 1) * M.g.d[d];
            og_params p{0, 5, 3, 0, 0, 1, 0};
            std::vector<double> tt(M.N);
            std::vector<int64_t> order(M.N);
            int64_t ninit = 0;
            const int64_t nacc = og_fmm_record(&M.g, M.vel.data(), src, &p, tt.data(), order.data(), nullptr, &ninit);
            if (nacc <= 0) {
                printf("oracle failed: %s\n", og_last_error());
                return 2;
            }
            std::vector<int64_t> init(order.begin(), order.begin() + ninit);
            for (int keep = 0; keep < 2 && do_filtered; keep++)
                for (int ver = 0; ver < 2; ver++)
                    for (int sd = 0; sd < nseeds; sd++) {
                        long passes = 0;
                        double mr = 0.0;
                        const long nd =
                            run_filtered(M, tt, init, keep != 0, ver != 0, 77u + 13u * (unsigned)sd, &passes, &mr);
                        total++;
                        if (nd != 0 && keep && ver) bad++;
                        if (nd != 0 && !(keep && ver)) unsafe_off++;
                        printf("%s src %d filtered+torn, key pools %s, %s, seed %d: %ld passes, %ld nodes differ from "
                               "the heap (max rel %.3e)\n", c.name, s, keep ? "always written" : "cascade-only",
                               ver ? "verified" : "unverified", sd, passes, nd, mr);
                    }
            for (int mode = 0; mode < 4 && do_orders; mode++)
                for (int sd = 0; sd < nseeds; sd++) {
                    long passes = 0;
                    double mr = 0.0;
                    const long nd = run(M, tt, init, mode, 1000u * (unsigned)mode + (unsigned)sd, &passes, &mr);
                    total++;
                    if (nd != 0) bad++;
                    printf("%s src %d %s seed %d: %ld passes, %ld nodes differ from the heap (max rel %.3e)\n", c.name,
                           s, names[mode], sd, passes, nd, mr);
                }
        }
    }
    printf("runs %ld, runs off the heap's answer %ld; without the final verification or with cascade-only key pools "
           "%ld\n", total, bad, unsafe_off);
    return bad != 0;
}
