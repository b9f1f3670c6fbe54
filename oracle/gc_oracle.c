/*
 * gc_oracle.c -- CPU oracle for the tiled push-relabel grid cut.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this file; it is the checker that tests/, the smoke test and the CPU
 * baseline leg of bench.py compare the HIP path against.
 *
 * Graph (the contract the HIP path implements):
 *   4-connected W x H grid, pixel p = y*W + x.
 *   cs[p] >= 0  source capacity,  ct[p] >= 0  sink capacity   (int32)
 *   right[p] = w(p, p+1)  (ignored in the last column)           (int32)
 *   down[p]  = w(p, p+W)  (ignored in the last row)              (int32)
 *   Pairwise arcs are symmetric: cap(p->q) = cap(q->p) = w(p,q).
 *
 * Two independent solvers live here:
 *
 *  1. oracle_pr_cut -- a sequential restatement of the reference's phase-1
 *     push-relabel in its "excess" formulation (e0 = cs - ct; pixels with
 *     e < 0 are sinks of height 0).  It follows:
 *       - driver order:       GraphCut.pyx:348-418 (startBfs, then push;
 *                             every `bfs`-th iteration a global relabel and
 *                             completion check, otherwise a local relabel)
 *       - push pass order:    GraphCut.pyx:242-323 (down, up, right, left)
 *       - chain push:         graphcut.cl:403-470 (push_down), 472-536
 *                             (push_up), 282-401 (push_right), 158-279
 *                             (push_left): 8-pixel chains, each carrying its
 *                             outflow into the next pixel; the carry that
 *                             leaves a chain is added to the next chain's
 *                             first pixel after the pass (graphcut.cl:457-469,
 *                             add_border graphcut.cl:114-156)
 *       - local relabel:      graphcut.cl:538-577 (Jacobi, capped at MAX)
 *       - global relabel:     graphcut.cl:582-1016 (distance to the sink set
 *                             {e < 0} over residual arcs).  Restated as an
 *                             exact BFS: the reference's 30-round cap on the
 *                             in-tile relaxation (graphcut.cl:928) can leave
 *                             distances unconverged, which is a defect, not
 *                             semantics.
 *       - completion:         graphcut.cl:20-60 (no pixel with e>0, h<MAX)
 *     Arithmetic is int64 (the reference uses float32).
 *
 *  2. oracle_bk_cut -- Boykov-Kolmogorov augmenting-path max-flow
 *     (Boykov & Kolmogorov, "An Experimental Comparison of Min-Cut/Max-Flow
 *     Algorithms for Energy Minimization in Vision", PAMI 2004), written from
 *     the paper: two search trees, growth / augmentation / adoption, with the
 *     timestamp + distance heuristic for choosing new parents.  It does not
 *     share any code with (1), so agreement of the two pins both.
 *
 * Outputs of both:
 *   F      = max-flow value (int64)
 *   mask   = minimal source set S_min: pixels reachable from the source in
 *            the final residual graph (1 = source / foreground side).
 *            For (1): reachable over residual pixel arcs from {e_final > 0}.
 *            For (2): reachable over residual pixel arcs from {tr > 0}.
 *   F      = sum_p min(cs,ct) + sum_p max(e0,0) - sum_p max(e_final,0)  (1)
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define DIR_UP 0
#define DIR_DOWN 1
#define DIR_LEFT 2
#define DIR_RIGHT 3

static inline int rev_dir(int d) { return d ^ 1; }

static inline int64_t nbr_of(int64_t p, int d, int W, int H)
{
    int64_t x = p % W, y = p / W;
    switch (d) {
    case DIR_UP:    return y > 0 ? p - W : -1;
    case DIR_DOWN:  return y < H - 1 ? p + W : -1;
    case DIR_LEFT:  return x > 0 ? p - 1 : -1;
    default:        return x < W - 1 ? p + 1 : -1;
    }
}

/* Residual capacities rc[4*p + d] for the symmetric 4-connected graph. */
static void build_caps(int W, int H, const int32_t *right, const int32_t *down,
                       int64_t *rc)
{
    int64_t n = (int64_t)W * H;
    memset(rc, 0, sizeof(int64_t) * 4 * n);
    for (int64_t y = 0; y < H; ++y)
        for (int64_t x = 0; x < W; ++x) {
            int64_t p = y * W + x;
            if (x < W - 1) {
                rc[4 * p + DIR_RIGHT] = right[p];
                rc[4 * (p + 1) + DIR_LEFT] = right[p];
            }
            if (y < H - 1) {
                rc[4 * p + DIR_DOWN] = down[p];
                rc[4 * (p + W) + DIR_UP] = down[p];
            }
        }
}

/* Pixels reachable over residual arcs from the seed set (seed[p] != 0). */
static void reach_from(int W, int H, const int64_t *rc, uint8_t *mask,
                       int64_t *queue)
{
    int64_t n = (int64_t)W * H, head = 0, tail = 0;
    for (int64_t p = 0; p < n; ++p)
        if (mask[p]) queue[tail++] = p;
    while (head < tail) {
        int64_t p = queue[head++];
        for (int d = 0; d < 4; ++d) {
            int64_t q = nbr_of(p, d, W, H);
            if (q >= 0 && !mask[q] && rc[4 * p + d] > 0) {
                mask[q] = 1;
                queue[tail++] = q;
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* 1. restatement of the reference's phase-1 push-relabel             */
/* ------------------------------------------------------------------ */

#define CHAIN 8 /* WAVE_LENGTH, graphcut.cl:9 */

/* Exact distance to the sink set {e < 0} over residual arcs; MAX if none.
 * Intent of init_bfs / bfs_intratile / bfs_intertile (graphcut.cl:582-1016). */
static void global_relabel(int W, int H, const int64_t *e, const int64_t *rc,
                           int32_t *h, int32_t maxh, int64_t *queue)
{
    int64_t n = (int64_t)W * H, head = 0, tail = 0;
    for (int64_t p = 0; p < n; ++p) {
        if (e[p] < 0) { h[p] = 0; queue[tail++] = p; }
        else h[p] = maxh;
    }
    while (head < tail) {
        int64_t q = queue[head++];
        for (int d = 0; d < 4; ++d) {
            int64_t p = nbr_of(q, d, W, H);
            /* arc p -> q has direction rev(d) seen from p */
            if (p >= 0 && h[p] == maxh && rc[4 * p + rev_dir(d)] > 0) {
                h[p] = h[q] + 1;
                queue[tail++] = p;
            }
        }
    }
}

/* One directional chain-push pass over the whole image.
 * dir: the direction flow moves (DIR_DOWN, DIR_UP, DIR_RIGHT, DIR_LEFT).
 * Chains are CHAIN pixels long, aligned at multiples of CHAIN along the
 * push axis (graphcut.cl:437-455 for down; the others are mirror images).
 * Heights are read, not written (graphcut.cl:435,441). */
static void push_pass(int W, int H, int dir, int64_t *e, int64_t *rc,
                      const int32_t *h, int32_t maxh, int64_t *carry)
{
    int vertical = (dir == DIR_DOWN || dir == DIR_UP);
    int len = vertical ? H : W;      /* extent along the push axis */
    int across = vertical ? W : H;   /* number of lines */
    int forward = (dir == DIR_DOWN || dir == DIR_RIGHT);
    int64_t n = (int64_t)W * H;
    memset(carry, 0, sizeof(int64_t) * n);

    for (int line = 0; line < across; ++line) {
        for (int c0 = 0; c0 < len; c0 += CHAIN) {
            int c1 = c0 + CHAIN < len ? c0 + CHAIN : len;
            int64_t flow = 0;
            int last_target = -1;
            for (int k = 0; k < c1 - c0; ++k) {
                int a = forward ? c0 + k : c1 - 1 - k;
                int64_t p = vertical ? (int64_t)a * W + line : (int64_t)line * W + a;
                int64_t q = nbr_of(p, dir, W, H);
                int64_t ef = e[p] + flow;
                if (q >= 0 && ef > 0 && h[p] < maxh && h[p] == h[q] + 1) {
                    flow = rc[4 * p + dir] < ef ? rc[4 * p + dir] : ef;
                } else {
                    flow = 0;
                }
                if (q >= 0) {
                    rc[4 * p + dir] -= flow;
                    rc[4 * q + rev_dir(dir)] += flow;
                }
                e[p] = ef - flow;
                last_target = (int)(q >= 0 ? q : -1);
                if (k == c1 - c0 - 1 && flow > 0 && last_target >= 0)
                    carry[last_target] += flow; /* leaves the chain */
            }
        }
    }
    for (int64_t p = 0; p < n; ++p) e[p] += carry[p];
}

/* Jacobi local relabel over the whole image (graphcut.cl:538-577). */
static void local_relabel(int W, int H, const int64_t *e, const int64_t *rc,
                          const int32_t *h, int32_t *h2, int32_t maxh)
{
    int64_t n = (int64_t)W * H;
    for (int64_t p = 0; p < n; ++p) {
        int32_t nh = h[p];
        if (e[p] > 0 && h[p] < maxh) {
            nh = maxh;
            for (int d = 0; d < 4; ++d) {
                int64_t q = nbr_of(p, d, W, H);
                if (q >= 0 && rc[4 * p + d] > 0 && h[q] + 1 < nh) nh = h[q] + 1;
            }
        }
        h2[p] = nh;
    }
}

static int has_active(int64_t n, const int64_t *e, const int32_t *h, int32_t maxh)
{
    for (int64_t p = 0; p < n; ++p)
        if (e[p] > 0 && h[p] < maxh) return 1;
    return 0;
}

/*
 * Returns 0 on success, -1 on allocation failure, -2 if max_iters was hit.
 * flow_out: max-flow value.  mask_out: W*H bytes, S_min.  iters_out: number
 * of push iterations performed (may be NULL).
 */
int oracle_pr_cut(int W, int H, const int32_t *cs, const int32_t *ct,
                  const int32_t *right, const int32_t *down, int bfs,
                  int64_t max_iters, int64_t *flow_out, uint8_t *mask_out,
                  int64_t *iters_out)
{
    int64_t n = (int64_t)W * H;
    int32_t maxh = (int32_t)n; /* MAX_HEIGHT = W*H, GraphCut.pyx:88 */
    if (bfs < 1) bfs = 1;
    int64_t *e = malloc(sizeof(int64_t) * n);
    int64_t *rc = malloc(sizeof(int64_t) * 4 * n);
    int64_t *tmp = malloc(sizeof(int64_t) * n);
    int32_t *h = malloc(sizeof(int32_t) * n);
    int32_t *h2 = malloc(sizeof(int32_t) * n);
    if (!e || !rc || !tmp || !h || !h2) {
        free(e); free(rc); free(tmp); free(h); free(h2);
        return -1;
    }
    int64_t base = 0, pos0 = 0;
    for (int64_t p = 0; p < n; ++p) {
        base += cs[p] < ct[p] ? cs[p] : ct[p];
        e[p] = (int64_t)cs[p] - ct[p];
        if (e[p] > 0) pos0 += e[p];
    }
    build_caps(W, H, right, down, rc);

    global_relabel(W, H, e, rc, h, maxh, tmp); /* startBfs, GraphCut.pyx:378 */
    int rcode = 0;
    int64_t iteration = 1;
    static const int order[4] = {DIR_DOWN, DIR_UP, DIR_RIGHT, DIR_LEFT};
    for (;;) {
        if (max_iters > 0 && iteration > max_iters) { rcode = -2; break; }
        for (int k = 0; k < 4; ++k) push_pass(W, H, order[k], e, rc, h, maxh, tmp);
        if (iteration % bfs == 0) {
            global_relabel(W, H, e, rc, h, maxh, tmp);
            if (!has_active(n, e, h, maxh)) break;
        } else {
            local_relabel(W, H, e, rc, h, h2, maxh);
            int32_t *t = h; h = h2; h2 = t;
        }
        ++iteration;
    }
    if (iters_out) *iters_out = iteration;

    int64_t posf = 0;
    for (int64_t p = 0; p < n; ++p) {
        if (e[p] > 0) posf += e[p];
        mask_out[p] = e[p] > 0;
    }
    reach_from(W, H, rc, mask_out, tmp);
    *flow_out = base + pos0 - posf;
    free(e); free(rc); free(tmp); free(h); free(h2);
    return rcode;
}

/* ------------------------------------------------------------------ */
/* 2. Boykov-Kolmogorov                                               */
/* ------------------------------------------------------------------ */

#define T_FREE 0
#define T_SRC 1
#define T_SNK 2
#define P_TERMINAL 4
#define P_ORPHAN 5
#define P_NONE 6
#define BK_INF 0x3fffffff

typedef struct {
    int W, H;
    int64_t n;
    int64_t *rc;  /* residual of p -> nbr(p,d) */
    int64_t *tr;  /* >0: residual s->p ; <0: residual p->t */
    uint8_t *tree;
    uint8_t *parent; /* direction from node to its parent, or P_* */
    int32_t *ts, *dist;
    uint8_t *inq;
    int64_t *aq; int64_t ahead, atail, acap; /* active FIFO (ring) */
    int64_t *oq; int64_t ohead, otail, ocap; /* orphan FIFO (ring) */
    int32_t time;
} bk_t;

static void bk_push_active(bk_t *g, int64_t p)
{
    if (g->inq[p]) return;
    g->inq[p] = 1;
    g->aq[g->atail] = p;
    g->atail = (g->atail + 1) % g->acap;
}

static int64_t bk_pop_active(bk_t *g)
{
    while (g->ahead != g->atail) {
        int64_t p = g->aq[g->ahead];
        g->ahead = (g->ahead + 1) % g->acap;
        g->inq[p] = 0;
        if (g->parent[p] != P_NONE) return p;
    }
    return -1;
}

static void bk_push_orphan(bk_t *g, int64_t p)
{
    g->parent[p] = P_ORPHAN;
    g->oq[g->otail] = p;
    g->otail = (g->otail + 1) % g->ocap;
}

/* Residual of the tree arc between p and its parent (flow direction:
 * parent -> p in the source tree, p -> parent in the sink tree). */
static inline int64_t *bk_tree_arc(bk_t *g, int64_t p, int64_t par)
{
    int d = g->parent[p];
    return g->tree[p] == T_SRC ? &g->rc[4 * par + rev_dir(d)] : &g->rc[4 * p + d];
}

static int64_t bk_augment(bk_t *g, int64_t sp, int db, int64_t tp)
{
    int64_t b = g->rc[4 * sp + db];
    int64_t p, par;
    for (p = sp; g->parent[p] != P_TERMINAL; p = par) {
        par = nbr_of(p, g->parent[p], g->W, g->H);
        int64_t c = *bk_tree_arc(g, p, par);
        if (c < b) b = c;
    }
    if (g->tr[p] < b) b = g->tr[p];
    for (p = tp; g->parent[p] != P_TERMINAL; p = par) {
        par = nbr_of(p, g->parent[p], g->W, g->H);
        int64_t c = *bk_tree_arc(g, p, par);
        if (c < b) b = c;
    }
    if (-g->tr[p] < b) b = -g->tr[p];

    g->rc[4 * sp + db] -= b;
    g->rc[4 * tp + rev_dir(db)] += b;
    for (p = sp; g->parent[p] != P_TERMINAL; p = par) {
        int d = g->parent[p];
        par = nbr_of(p, d, g->W, g->H);
        g->rc[4 * par + rev_dir(d)] -= b;
        g->rc[4 * p + d] += b;
        if (g->rc[4 * par + rev_dir(d)] == 0) bk_push_orphan(g, p);
    }
    g->tr[p] -= b;
    if (g->tr[p] == 0) bk_push_orphan(g, p);
    for (p = tp; g->parent[p] != P_TERMINAL; p = par) {
        int d = g->parent[p];
        par = nbr_of(p, d, g->W, g->H);
        g->rc[4 * p + d] -= b;
        g->rc[4 * par + rev_dir(d)] += b;
        if (g->rc[4 * p + d] == 0) bk_push_orphan(g, p);
    }
    g->tr[p] += b;
    if (g->tr[p] == 0) bk_push_orphan(g, p);
    return b;
}

static void bk_adopt(bk_t *g)
{
    while (g->ohead != g->otail) {
        int64_t p = g->oq[g->ohead];
        g->ohead = (g->ohead + 1) % g->ocap;
        int tr = g->tree[p];
        int best = -1;
        int32_t dmin = BK_INF;
        for (int d = 0; d < 4; ++d) {
            int64_t q = nbr_of(p, d, g->W, g->H);
            if (q < 0 || g->tree[q] != tr || g->parent[q] == P_NONE) continue;
            int64_t res = tr == T_SRC ? g->rc[4 * q + rev_dir(d)] : g->rc[4 * p + d];
            if (res <= 0) continue;
            /* trace q to its origin */
            int64_t j = q;
            int32_t dd = 0;
            for (;;) {
                if (g->ts[j] == g->time) { dd += g->dist[j]; break; }
                int a = g->parent[j];
                dd++;
                if (a == P_TERMINAL) { g->ts[j] = g->time; g->dist[j] = 1; break; }
                if (a == P_ORPHAN) { dd = BK_INF; break; }
                j = nbr_of(j, a, g->W, g->H);
            }
            if (dd < BK_INF) {
                if (dd < dmin) { dmin = dd; best = d; }
                for (j = q; g->ts[j] != g->time; j = nbr_of(j, g->parent[j], g->W, g->H)) {
                    g->ts[j] = g->time;
                    g->dist[j] = dd--;
                }
            }
        }
        if (best >= 0) {
            g->parent[p] = (uint8_t)best;
            g->ts[p] = g->time;
            g->dist[p] = dmin + 1;
            continue;
        }
        /* no parent: p becomes free */
        for (int d = 0; d < 4; ++d) {
            int64_t q = nbr_of(p, d, g->W, g->H);
            if (q < 0 || g->tree[q] != tr || g->parent[q] == P_NONE) continue;
            int64_t res = tr == T_SRC ? g->rc[4 * q + rev_dir(d)] : g->rc[4 * p + d];
            if (res > 0) bk_push_active(g, q);
            int a = g->parent[q];
            if (a < 4 && nbr_of(q, a, g->W, g->H) == p) bk_push_orphan(g, q);
        }
        g->parent[p] = P_NONE;
        g->tree[p] = T_FREE;
    }
}

int oracle_bk_cut(int W, int H, const int32_t *cs, const int32_t *ct,
                  const int32_t *right, const int32_t *down,
                  int64_t *flow_out, uint8_t *mask_out)
{
    bk_t g;
    memset(&g, 0, sizeof g);
    g.W = W; g.H = H; g.n = (int64_t)W * H;
    int64_t n = g.n;
    g.rc = malloc(sizeof(int64_t) * 4 * n);
    g.tr = malloc(sizeof(int64_t) * n);
    g.tree = calloc(n, 1);
    g.parent = malloc(n);
    g.ts = calloc(n, sizeof(int32_t));
    g.dist = calloc(n, sizeof(int32_t));
    g.inq = calloc(n, 1);
    g.acap = n + 1; g.ocap = n + 1;
    g.aq = malloc(sizeof(int64_t) * g.acap);
    g.oq = malloc(sizeof(int64_t) * g.ocap);
    if (!g.rc || !g.tr || !g.tree || !g.parent || !g.ts || !g.dist || !g.inq ||
        !g.aq || !g.oq) {
        free(g.rc); free(g.tr); free(g.tree); free(g.parent); free(g.ts);
        free(g.dist); free(g.inq); free(g.aq); free(g.oq);
        return -1;
    }
    build_caps(W, H, right, down, g.rc);
    int64_t flow = 0;
    for (int64_t p = 0; p < n; ++p) {
        flow += cs[p] < ct[p] ? cs[p] : ct[p];
        g.tr[p] = (int64_t)cs[p] - ct[p];
        if (g.tr[p] > 0) {
            g.tree[p] = T_SRC; g.parent[p] = P_TERMINAL; g.dist[p] = 1;
            bk_push_active(&g, p);
        } else if (g.tr[p] < 0) {
            g.tree[p] = T_SNK; g.parent[p] = P_TERMINAL; g.dist[p] = 1;
            bk_push_active(&g, p);
        } else {
            g.tree[p] = T_FREE; g.parent[p] = P_NONE;
        }
    }

    int64_t cur = -1;
    for (;;) {
        if (cur < 0 || g.parent[cur] == P_NONE) {
            cur = bk_pop_active(&g);
            if (cur < 0) break;
        }
        int64_t sp = -1, tp = -1;
        int db = -1;
        if (g.tree[cur] == T_SRC) {
            for (int d = 0; d < 4 && sp < 0; ++d) {
                int64_t q = nbr_of(cur, d, W, H);
                if (q < 0 || g.rc[4 * cur + d] <= 0) continue;
                if (g.parent[q] == P_NONE) {
                    g.tree[q] = T_SRC; g.parent[q] = (uint8_t)rev_dir(d);
                    g.ts[q] = g.ts[cur]; g.dist[q] = g.dist[cur] + 1;
                    bk_push_active(&g, q);
                } else if (g.tree[q] == T_SNK) {
                    sp = cur; tp = q; db = d;
                } else if (g.ts[q] <= g.ts[cur] && g.dist[q] > g.dist[cur]) {
                    g.parent[q] = (uint8_t)rev_dir(d);
                    g.ts[q] = g.ts[cur]; g.dist[q] = g.dist[cur] + 1;
                }
            }
        } else {
            for (int d = 0; d < 4 && sp < 0; ++d) {
                int64_t q = nbr_of(cur, d, W, H);
                if (q < 0 || g.rc[4 * q + rev_dir(d)] <= 0) continue;
                if (g.parent[q] == P_NONE) {
                    g.tree[q] = T_SNK; g.parent[q] = (uint8_t)rev_dir(d);
                    g.ts[q] = g.ts[cur]; g.dist[q] = g.dist[cur] + 1;
                    bk_push_active(&g, q);
                } else if (g.tree[q] == T_SRC) {
                    sp = q; tp = cur; db = rev_dir(d);
                } else if (g.ts[q] <= g.ts[cur] && g.dist[q] > g.dist[cur]) {
                    g.parent[q] = (uint8_t)rev_dir(d);
                    g.ts[q] = g.ts[cur]; g.dist[q] = g.dist[cur] + 1;
                }
            }
        }
        g.time++;
        if (sp >= 0) {
            flow += bk_augment(&g, sp, db, tp);
            bk_adopt(&g);
            /* keep growing from cur while it stays in a tree */
        } else {
            cur = -1;
        }
    }

    for (int64_t p = 0; p < n; ++p) mask_out[p] = g.tr[p] > 0;
    reach_from(W, H, g.rc, mask_out, g.aq);
    *flow_out = flow;
    free(g.rc); free(g.tr); free(g.tree); free(g.parent); free(g.ts);
    free(g.dist); free(g.inq); free(g.aq); free(g.oq);
    return 0;
}

/* Cut value of a labelling (1 = source side): sum of ct over source-side
 * pixels + cs over sink-side pixels + w(p,q) over arcs p in S, q not in S. */
int64_t oracle_cut_value(int W, int H, const int32_t *cs, const int32_t *ct,
                         const int32_t *right, const int32_t *down,
                         const uint8_t *mask)
{
    int64_t v = 0;
    for (int64_t y = 0; y < H; ++y)
        for (int64_t x = 0; x < W; ++x) {
            int64_t p = y * W + x;
            v += mask[p] ? ct[p] : cs[p];
            if (x < W - 1 && mask[p] != mask[p + 1]) v += right[p];
            if (y < H - 1 && mask[p] != mask[p + W]) v += down[p];
        }
    return v;
}
