/*
 * oracle.c -- CPU restatement of the cc3d labelling path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * bench's cpu_baseline leg may load this library, and only as the checker.
 * The product path (the HIP library) never links, loads or calls it.
 *
 * What it restates (reference files cited per function below):
 *   - cc3d/cc3d.hpp:287-315      estimate_provisional_label_count
 *   - cc3d/cc3d.hpp:344-1421     multilabel first passes (26/18/6, 2D 4/8)
 *   - cc3d/cc3d_binary.hpp       binary first passes incl. 2x2 block 2D-8
 *   - cc3d/cc3d_continuous.hpp   continuous (|a-b| <= delta) first passes
 *   - cc3d/cc3d.hpp:239-285      relabel (first-appearance numbering)
 *
 * The reference's result is determined by the set of union-find links its
 * decision trees record (copies + unifies).  The restatement computes, per
 * voxel, a link set whose transitive closure equals the reference's, then
 * numbers components by the rank of their first voxel in memory order
 * (x fastest): a component's first voxel always receives a fresh provisional
 * label and relabel() assigns final labels in provisional order.
 *
 *   canonical  (multilabel, binary, continuous 3D / 2D-4):
 *              p links to every backward neighbour q with
 *                multilabel : v[q] == v[p]           (T's ==, v[p] != 0)
 *                binary     : v[q] != 0
 *                continuous : v[q] != 0 && match(v[p], v[q], delta)
 *              plus, for continuous 26-connectivity, the exact-equality link
 *              to the voxel below (cc3d/cc3d_continuous.hpp:147-150).
 *   quirks     Q1 binary 2D-8 numbers by first 2x2 block
 *                 (cc3d/cc3d_binary.hpp:955-1263)
 *              Q2 binary 2D-4: row 0 links on value equality; later rows
 *                 skip the left link when the up and up-left pixels are set
 *                 (cc3d/cc3d_binary.hpp:822-935)
 *              Q3 continuous 2D-8 links only the pixel above when it is equal
 *                 or when both lie within delta of the global min/max
 *                 (cc3d/cc3d_continuous.hpp:298-349)
 *              Q6 binary 2D-8 leaves provisional block labels in background
 *                 pixels outside the rewritten runs
 *                 (cc3d/cc3d_binary.hpp:1233-1260)
 * P, the provisional label count, is the number of voxels (blocks for Q1)
 * whose link set is empty; the reference throws "maximum length exception"
 * when P reaches its union-find length (cc3d/cc3d.hpp:107-111).
 *
 * Pinning: tests/test_oracle_*.py check every function here against the
 * compiled reference (oracle/_ref), the reference's known-answer tests
 * (tests/golden/), scipy.ndimage.label and brute-force BFS.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum { DT_U8 = 0, DT_U16 = 1, DT_U32 = 2, DT_U64 = 3, DT_F32 = 4, DT_F64 = 5 };
enum { M_MULTI = 0, M_BINARY = 1, M_CONT = 2 };

/* Deliberate mistakes, for tests that show the checks catch them:
 * 1 drop the 26-connected corner neighbours, 2 number components by their
 * last voxel, 3 ignore the binary 2D-4 row-0 equality rule, 4 ignore the
 * continuous 2D-8 range shortcut, 5 skip the stale binary 2D-8 labels. */
static int g_mutation = 0;
void oracle_set_mutation(int m) { g_mutation = m; }

/* ---------------- union-find (plain, sequential) ---------------- */
static uint32_t uf_find(uint32_t* P, uint32_t a) {
  while (P[a] != a) {
    P[a] = P[P[a]];
    a = P[a];
  }
  return a;
}
static void uf_union(uint32_t* P, uint32_t a, uint32_t b) {
  a = uf_find(P, a);
  b = uf_find(P, b);
  if (a == b) return;
  if (a < b) P[b] = a; else P[a] = b;
}

/* ---------------- typed value helpers ---------------- */
typedef struct {
  const void* p;
  int dt;
} vol_t;

static inline int is_fg(const vol_t* v, int64_t i) {
  switch (v->dt) {
    case DT_U8: return ((const uint8_t*)v->p)[i] != 0;
    case DT_U16: return ((const uint16_t*)v->p)[i] != 0;
    case DT_U32: return ((const uint32_t*)v->p)[i] != 0;
    case DT_U64: return ((const uint64_t*)v->p)[i] != 0;
    case DT_F32: return ((const float*)v->p)[i] != 0.0f;
    default: return ((const double*)v->p)[i] != 0.0;
  }
}
static inline int is_eq(const vol_t* v, int64_t i, int64_t j) {
  switch (v->dt) {
    case DT_U8: return ((const uint8_t*)v->p)[i] == ((const uint8_t*)v->p)[j];
    case DT_U16: return ((const uint16_t*)v->p)[i] == ((const uint16_t*)v->p)[j];
    case DT_U32: return ((const uint32_t*)v->p)[i] == ((const uint32_t*)v->p)[j];
    case DT_U64: return ((const uint64_t*)v->p)[i] == ((const uint64_t*)v->p)[j];
    case DT_F32: return ((const float*)v->p)[i] == ((const float*)v->p)[j];
    default: return ((const double*)v->p)[i] == ((const double*)v->p)[j];
  }
}
/* cc3d/cc3d_continuous.hpp:79-88: unsigned -> (a>b ? a-b : b-a) <= d;
 * float/double -> fabs(a - b) <= d evaluated in T. */
static inline int is_match(const vol_t* v, int64_t i, int64_t j, uint64_t du,
                           double df) {
  switch (v->dt) {
    case DT_U8: {
      uint8_t a = ((const uint8_t*)v->p)[i], b = ((const uint8_t*)v->p)[j];
      uint8_t d = (uint8_t)(a > b ? a - b : b - a);
      return d <= (uint8_t)du;
    }
    case DT_U16: {
      uint16_t a = ((const uint16_t*)v->p)[i], b = ((const uint16_t*)v->p)[j];
      uint16_t d = (uint16_t)(a > b ? a - b : b - a);
      return d <= (uint16_t)du;
    }
    case DT_U32: {
      uint32_t a = ((const uint32_t*)v->p)[i], b = ((const uint32_t*)v->p)[j];
      return (uint32_t)(a > b ? a - b : b - a) <= (uint32_t)du;
    }
    case DT_U64: {
      uint64_t a = ((const uint64_t*)v->p)[i], b = ((const uint64_t*)v->p)[j];
      return (a > b ? a - b : b - a) <= du;
    }
    case DT_F32: {
      volatile float a = ((const float*)v->p)[i], b = ((const float*)v->p)[j];
      volatile float d = a - b;
      return fabsf(d) <= (float)df;
    }
    default: {
      volatile double a = ((const double*)v->p)[i], b = ((const double*)v->p)[j];
      volatile double d = a - b;
      return fabs(d) <= df;
    }
  }
}

/* ---------------- estimate_provisional_label_count ----------------
 * cc3d/cc3d.hpp:287-315: per row, count the first voxel if nonzero plus every
 * x where v[x] != v[x-1] and v[x] != 0; track first/last rows with a count. */
uint64_t oracle_epl(const void* in, int dt, int64_t sx, int64_t voxels,
                    int64_t* first_row, int64_t* last_row) {
  vol_t v = {in, dt};
  int64_t f = -1, l = -1;
  uint64_t count = 0;
  int64_t row = 0;
  for (int64_t loc = 0; loc < voxels; loc += sx, row++) {
    uint64_t rc = is_fg(&v, loc);
    for (int64_t x = 1; x < sx; x++)
      rc += (!is_eq(&v, loc + x, loc + x - 1) && is_fg(&v, loc + x));
    count += rc;
    if (rc) {
      if (f == -1) f = row;
      l = row;
    }
  }
  *first_row = f;
  *last_row = l;
  return count;
}

/* ---------------- neighbourhoods ---------------- */
typedef struct {
  int dx, dy, dz;
} off_t3;

static int backward_offsets(int conn, off_t3* o) {
  int n = 0;
  if (conn == 4 || conn == 8) {
    o[n++] = (off_t3){-1, 0, 0};
    o[n++] = (off_t3){0, -1, 0};
    if (conn == 8) {
      o[n++] = (off_t3){-1, -1, 0};
      o[n++] = (off_t3){1, -1, 0};
    }
    return n;
  }
  for (int dz = -1; dz <= 0; dz++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) {
        if (dz == 0 && (dy > 0 || (dy == 0 && dx >= 0))) continue;
        int m = abs(dx) + abs(dy) + abs(dz);
        if (conn == 6 && m > 1) continue;
        if (conn == 18 && m > 2) continue;
        o[n++] = (off_t3){dx, dy, dz};
      }
  return n;
}

/* ---------------- global min/max for Q3 ----------------
 * cc3d/cc3d_continuous.hpp:298-303 folds std::max/std::min over the raw
 * array starting from element 0; std::max(a,b) = (a < b) ? b : a. */
static void q3_minmax(const vol_t* v, int64_t voxels, double* gmin_f,
                      double* gmax_f, uint64_t* gmin_u, uint64_t* gmax_u) {
  if (v->dt == DT_F32) {
    const float* a = (const float*)v->p;
    float mx = a[0], mn = a[0];
    for (int64_t i = 1; i < voxels; i++) {
      mx = (a[i] < mx) ? mx : a[i];
      mn = (mn < a[i]) ? mn : a[i];
    }
    *gmin_f = mn;
    *gmax_f = mx;
  } else if (v->dt == DT_F64) {
    const double* a = (const double*)v->p;
    double mx = a[0], mn = a[0];
    for (int64_t i = 1; i < voxels; i++) {
      mx = (a[i] < mx) ? mx : a[i];
      mn = (mn < a[i]) ? mn : a[i];
    }
    *gmin_f = mn;
    *gmax_f = mx;
  } else {
    uint64_t mx = 0, mn = 0;
    for (int64_t i = 0; i < voxels; i++) {
      uint64_t x;
      switch (v->dt) {
        case DT_U8: x = ((const uint8_t*)v->p)[i]; break;
        case DT_U16: x = ((const uint16_t*)v->p)[i]; break;
        case DT_U32: x = ((const uint32_t*)v->p)[i]; break;
        default: x = ((const uint64_t*)v->p)[i]; break;
      }
      if (i == 0) mx = mn = x;
      if (x > mx) mx = x;
      if (x < mn) mn = x;
    }
    *gmin_u = mn;
    *gmax_u = mx;
  }
}

/* Q3 shortcut predicate: min(cur,B) - gmin <= delta && gmax - max(cur,B) <= delta
 * in T arithmetic (cc3d/cc3d_continuous.hpp:341-349). */
static int q3_span(const vol_t* v, int64_t p, int64_t b, uint64_t du, double df,
                   double gmin_f, double gmax_f, uint64_t gmin_u, uint64_t gmax_u) {
  if (v->dt == DT_F32) {
    float c = ((const float*)v->p)[p], q = ((const float*)v->p)[b];
    float lo = (q < c) ? q : c;  /* std::min(cur, B) */
    float hi = (c < q) ? q : c;  /* std::max(cur, B) */
    volatile float d1 = lo - (float)gmin_f;
    volatile float d2 = (float)gmax_f - hi;
    return d1 <= (float)df && d2 <= (float)df;
  } else if (v->dt == DT_F64) {
    double c = ((const double*)v->p)[p], q = ((const double*)v->p)[b];
    double lo = (q < c) ? q : c;
    double hi = (c < q) ? q : c;
    volatile double d1 = lo - gmin_f;
    volatile double d2 = gmax_f - hi;
    return d1 <= df && d2 <= df;
  } else {
    uint64_t c, q, mask;
    switch (v->dt) {
      case DT_U8: c = ((const uint8_t*)v->p)[p]; q = ((const uint8_t*)v->p)[b]; mask = 0xFFu; break;
      case DT_U16: c = ((const uint16_t*)v->p)[p]; q = ((const uint16_t*)v->p)[b]; mask = 0xFFFFu; break;
      case DT_U32: c = ((const uint32_t*)v->p)[p]; q = ((const uint32_t*)v->p)[b]; mask = 0xFFFFFFFFu; break;
      default: c = ((const uint64_t*)v->p)[p]; q = ((const uint64_t*)v->p)[b]; mask = ~0ull; break;
    }
    uint64_t lo = q < c ? q : c, hi = c < q ? q : c;
    /* arithmetic is in T after integer promotion; results are non-negative */
    return ((lo - gmin_u) & mask) <= (du & mask) && ((gmax_u - hi) & mask) <= (du & mask);
  }
}

/* ---------------- Q1 / Q6 helpers (binary 2D-8, 2x2 blocks) ---------------- */

/* Source block chosen by the block decision tree of
 * cc3d/cc3d_binary.hpp:1029-1161 for the block at pixel (x, y) (x, y even),
 * or -1 when the block takes a new provisional label, or -2 when the block
 * holds no foreground.  Only needed to reproduce provisional labels for Q6
 * and to count provisional labels for the length check. */
static int64_t q1_source(const vol_t* v, int64_t x, int64_t y, int64_t sx,
                         int64_t sy) {
  const int64_t osx = (sx + 1) >> 1;
  const int64_t ob = (x >> 1) + osx * (y >> 1);
#define PX(xx, yy) is_fg(v, (xx) + sx * (yy))
  const int Y = PX(x, y);
  const int Z = (x < sx - 1) && PX(x + 1, y);
  const int X = (y < sy - 1) && PX(x, y + 1);
  const int W = (x < sx - 1) && (y < sy - 1) && PX(x + 1, y + 1);
  const int64_t up = ob - osx, left = ob - 1, upleft = ob - 1 - osx,
                upright = ob + 1 - osx;
  if (Y) {
    if (y > 0 && PX(x, y - 1)) return up;                         /* D */
    if (y > 0 && x < sx - 1 && PX(x + 1, y - 1)) return up;       /* E */
    if (x > 0 && PX(x - 1, y)) return left;                       /* B */
    if (x > 0 && y > 0 && PX(x - 1, y - 1)) return upleft;        /* C */
    if (x > 0 && y < sy - 1 && PX(x - 1, y + 1)) return left;     /* A */
    if (x < sx - 2 && y > 0 && Z && PX(x + 2, y - 1)) return upright; /* Z&F */
    return -1;
  }
  if (Z) {
    if (y > 0 && PX(x + 1, y - 1)) return up;                     /* E */
    if (y > 0 && PX(x, y - 1)) return up;                         /* D */
    if (x < sx - 2 && y > 0 && PX(x + 2, y - 1)) return upright;  /* F */
    if (X) {
      if (x > 0 && PX(x - 1, y)) return left;                     /* B */
      if (x > 0 && PX(x - 1, y + 1)) return left;                 /* A */
      return -1;
    }
    return -1;
  }
  if (X) {
    if (x > 0 && PX(x - 1, y)) return left;
    if (x > 0 && PX(x - 1, y + 1)) return left;
    return -1;
  }
  if (W) return -1;
  return -2;
#undef PX
}

/* compute_foreground_index (cc3d/cc3d.hpp:317-342) for one row */
static void row_run(const vol_t* v, int64_t row, int64_t sx, uint32_t* s,
                    uint32_t* e) {
  uint32_t a = 0, b = 0;
  int64_t loc = row * sx;
  for (int64_t x = 0; x < sx; x++)
    if (is_fg(v, loc + x)) {
      a = (uint32_t)x;
      break;
    }
  for (int64_t x = sx - 1; x >= (int64_t)a; x--)
    if (is_fg(v, loc + x)) {
      b = (uint32_t)(x + 1);
      break;
    }
  *s = a;
  *e = b;
}

/* ---------------- main entry ---------------- */

/*
 * oracle_label: labels `in` (already viewed as u8/u16/u32/u64/f32/f64, x
 * fastest, shape sx*sy*sz) the way cc3d::connected_components3d<T,U>
 * (cc3d/cc3d_continuous.hpp:394-455) does.  Labels are written as uint64 to
 * out[]; *N gets the component count, *P the provisional label count.
 * mode: 0 multilabel (delta == 0), 1 binary, 2 continuous.
 * Returns 0, or -1 on bad arguments / allocation failure.
 */
int oracle_label(const void* in, int dt, int64_t sx, int64_t sy, int64_t sz,
                 int connectivity, int mode, uint64_t delta_u, double delta_f,
                 int periodic, uint64_t* out, uint64_t* N, uint64_t* P) {
  const int64_t sxy = sx * sy, V = sxy * sz;
  vol_t v = {in, dt};
  *N = 0;
  *P = 0;
  if (V <= 0) return 0;
  if (V >= 0xFFFFFFFFll) return -1;
  if ((connectivity == 4 || connectivity == 8) && sz != 1) return -1;

  const int q1 = (mode == M_BINARY && connectivity == 8);
  const int q2 = (mode == M_BINARY && connectivity == 4);
  const int q3 = (mode == M_CONT && connectivity == 8);

  uint32_t* par = (uint32_t*)malloc(sizeof(uint32_t) * V);
  if (!par) return -1;
  for (int64_t i = 0; i < V; i++) par[i] = (uint32_t)i;

  off_t3 offs[13];
  const int noff = backward_offsets(connectivity, offs);

  double gmin_f = 0, gmax_f = 0;
  uint64_t gmin_u = 0, gmax_u = 0;
  if (q3) q3_minmax(&v, V, &gmin_f, &gmax_f, &gmin_u, &gmax_u);

  uint64_t prov = 0;
  for (int64_t z = 0; z < sz; z++)
    for (int64_t y = 0; y < sy; y++)
      for (int64_t x = 0; x < sx; x++) {
        const int64_t p = x + sx * (y + sy * z);
        if (!is_fg(&v, p)) continue;
        int nlinks = 0;
        if (q2) {
          /* cc3d/cc3d_binary.hpp:822-935 */
          if (y == 0) {
            const int lk = x > 0 && (g_mutation == 3 ? is_fg(&v, p - 1) : is_eq(&v, p, p - 1));
            if (lk) { uf_union(par, p, p - 1); nlinks++; }
          } else {
            const int B = x > 0 && is_fg(&v, p - 1);
            const int C = is_fg(&v, p - sx);
            const int D = x > 0 && is_fg(&v, p - 1 - sx);
            if (C) { uf_union(par, p, p - sx); nlinks++; }
            if (B && !(C && D)) { uf_union(par, p, p - 1); nlinks++; }
          }
        } else if (q3) {
          /* cc3d/cc3d_continuous.hpp:335-384 */
          if (y > 0 && is_eq(&v, p, p - sx)) {
            uf_union(par, p, p - sx);
            nlinks++;
          } else if (g_mutation != 4 && y > 0 && is_fg(&v, p - sx) &&
                     q3_span(&v, p, p - sx, delta_u, delta_f, gmin_f, gmax_f, gmin_u, gmax_u)) {
            uf_union(par, p, p - sx);
            nlinks++;
          } else {
            const int64_t cand[4] = {y > 0 ? p - sx : -1, (x > 0 && y > 0) ? p - 1 - sx : -1,
                                     (x < sx - 1 && y > 0) ? p + 1 - sx : -1, x > 0 ? p - 1 : -1};
            for (int k = 0; k < 4; k++) {
              if (cand[k] < 0) continue;
              if (is_fg(&v, cand[k]) && is_match(&v, p, cand[k], delta_u, delta_f)) {
                uf_union(par, p, cand[k]);
                nlinks++;
              }
            }
          }
        } else {
          for (int k = 0; k < noff; k++) {
            const int64_t qx = x + offs[k].dx, qy = y + offs[k].dy, qz = z + offs[k].dz;
            if (qx < 0 || qx >= sx || qy < 0 || qy >= sy || qz < 0) continue;
            if (g_mutation == 1 && offs[k].dx && offs[k].dy && offs[k].dz) continue;
            const int64_t q = qx + sx * (qy + sy * qz);
            int lk;
            if (mode == M_MULTI) lk = is_eq(&v, p, q);
            else if (mode == M_BINARY) lk = is_fg(&v, q);
            else {
              lk = is_fg(&v, q) && is_match(&v, p, q, delta_u, delta_f);
              /* cc3d/cc3d_continuous.hpp:147-150 */
              if (connectivity == 26 && offs[k].dx == 0 && offs[k].dy == 0 && is_eq(&v, p, q))
                lk = 1;
            }
            if (lk) {
              uf_union(par, p, q);
              nlinks++;
            }
          }
        }
        if (!q1 && nlinks == 0) prov++;
      }

  /* periodic boundary links: cc3d/cc3d.hpp:1048-1073, :1265-1277,
   * :1377-1418; cc3d/cc3d_binary.hpp:733-758, :938-950, :1163-1210 */
  if (periodic && mode != M_CONT) {
#define PLINK(a, b)                                                         \
  do {                                                                      \
    int64_t _a = (a), _b = (b);                                             \
    if (is_fg(&v, _a) && (mode == M_BINARY ? is_fg(&v, _b) : is_eq(&v, _a, _b))) \
      uf_union(par, _a, _b);                                                \
  } while (0)
    if (connectivity == 6) {
      for (int64_t z = 0; z < sz; z++)
        for (int64_t y = 0; y < sy; y++) PLINK(sx * (y + sy * z), sx * (y + sy * z) + sx - 1);
      for (int64_t z = 0; z < sz; z++)
        for (int64_t x = 0; x < sx; x++) PLINK(x + sxy * z, x + sxy * z + sx * (sy - 1));
      for (int64_t y = 0; y < sy; y++)
        for (int64_t x = 0; x < sx; x++) PLINK(x + sx * y, x + sx * y + sxy * (sz - 1));
    } else if (connectivity == 4) {
      for (int64_t x = 0; x < sx; x++) PLINK(x, x + sx * (sy - 1));
      for (int64_t y = 0; y < sy; y++) PLINK(sx * y, sx * y + sx - 1);
    } else if (connectivity == 8) {
      for (int64_t x = 0; x < sx; x++) {
        if (x > 0) PLINK(x, x - 1 + sx * (sy - 1));
        PLINK(x, x + sx * (sy - 1));
        if (x < sx - 1) PLINK(x, x + 1 + sx * (sy - 1));
      }
      PLINK(0, V - 1);
      PLINK(sx - 1, sx * (sy - 1));
      for (int64_t y = 0; y < sy; y++) {
        PLINK(sx * y, sx * y + sx - 1);
        if (y > 0) PLINK(sx * y, sx * y + sx - 1 - sx);
        if (y < sy - 1) PLINK(sx * y, sx * y + sx - 1 + sx);
      }
    }
#undef PLINK
  }

  /* numbering key: voxel index, or for Q1 the 2x2 block key */
  const int64_t osx = (sx + 1) >> 1, osy = (sy + 1) >> 1;
  const int64_t K = q1 ? 4 * osx * osy : V;
  uint32_t* minkey = (uint32_t*)malloc(sizeof(uint32_t) * V);
  uint8_t* used = (uint8_t*)calloc(K, 1);
  uint64_t* rank = (uint64_t*)malloc(sizeof(uint64_t) * K);
  if (!minkey || !used || !rank) {
    free(par); free(minkey); free(used); free(rank);
    return -1;
  }
  for (int64_t i = 0; i < V; i++) minkey[i] = 0xFFFFFFFFu;
  for (int64_t p = 0; p < V; p++) {
    if (!is_fg(&v, p)) continue;
    uint32_t key;
    if (q1) {
      const int64_t x = p % sx, y = p / sx;
      key = (uint32_t)((((y >> 1) * osx + (x >> 1)) << 2) | ((y & 1) << 1) | (x & 1));
    } else {
      key = (uint32_t)p;
    }
    uint32_t r = uf_find(par, (uint32_t)p);
    if (g_mutation == 2) key = (uint32_t)(K - 1 - key);
    if (key < minkey[r]) minkey[r] = key;
  }
  for (int64_t p = 0; p < V; p++)
    if (is_fg(&v, p) && par[p] == (uint32_t)p) used[minkey[p]] = 1;
  uint64_t n = 0;
  for (int64_t k = 0; k < K; k++) {
    if (used[k]) n++;
    rank[k] = n;
  }
  for (int64_t p = 0; p < V; p++)
    out[p] = is_fg(&v, p) ? rank[minkey[uf_find(par, (uint32_t)p)]] : 0;
  *N = n;

  if (q1 && g_mutation != 5) {
    /* provisional block labels in block raster order (for P and Q6) */
    const int64_t nb = osx * osy;
    uint64_t* bl = (uint64_t*)calloc(nb, sizeof(uint64_t));
    if (!bl) {
      free(par); free(minkey); free(used); free(rank);
      return -1;
    }
    for (int64_t y = 0; y < sy; y += 2)
      for (int64_t x = 0; x < sx; x += 2) {
        const int64_t ob = (x >> 1) + osx * (y >> 1);
        const int64_t src = q1_source(&v, x, y, sx, sy);
        if (src == -2) continue;
        if (src == -1) bl[ob] = ++prov;
        else bl[ob] = bl[src];
      }
    /* Q6: cc3d/cc3d_binary.hpp:1233-1260 rewrites rows with y*sx < V/4
     * entirely, later rows only on [xstart, xend) of rows y and y+1; the
     * first osx*osy entries of out held the block labels. */
    const int64_t target = V >> 2;
    for (int64_t y = 0; y < sy; y++) {
      if (y * sx < target) continue;
      const int64_t ymax = (y < sy - 1) ? y + 1 : y;
      uint32_t s0, e0, s1, e1;
      row_run(&v, y, sx, &s0, &e0);
      row_run(&v, ymax, sx, &s1, &e1);
      const int64_t xs = (int64_t)(s0 < s1 ? s0 : s1) & ~(int64_t)1;
      const int64_t xe = (int64_t)(e0 > e1 ? e0 : e1);
      for (int64_t x = 0; x < sx; x++) {
        if (x >= xs && x < xe) continue;
        const int64_t loc = x + sx * y;
        if (loc < nb) out[loc] = bl[loc];
      }
    }
    free(bl);
  }
  *P = prov;

  free(par);
  free(minkey);
  free(used);
  free(rank);
  return 0;
}

/*
 * oracle_statistics: cc3d.statistics restated (cc3d/fastcc3d.pyx:681-765,
 * _statistics_helper2d :771-844, _statistics_helper3d :850-938).  labels is
 * the unsigned view (dt 0-3) of an array of numpy shape (s0, s1, s2) (s2 = 1
 * for 2D) in F (f_order = 1, axis 0 fastest) or C order.  Per label 0..N:
 * counts (uint32, wrapping like the reference's uint32 array), per numpy
 * axis the min and max coordinate (bbox[l*6 + 2a], bbox[l*6 + 2a + 1]; mins
 * start at UINT32_MAX and maxes at 0 for absent labels) and the sum of
 * coordinates (sums[l*3 + a]; the reference accumulates them in float64,
 * where they are exact integers below 2^53).  Returns 1 if a label exceeds
 * N.
 */
int oracle_statistics(const void* labels, int dt, int64_t s0, int64_t s1, int64_t s2, int f_order,
                      uint64_t N, uint32_t* counts, uint32_t* bbox, uint64_t* sums) {
  for (uint64_t l = 0; l <= N; l++) {
    counts[l] = 0;
    for (int a = 0; a < 3; a++) {
      bbox[l * 6 + 2 * a] = 0xFFFFFFFFu;
      bbox[l * 6 + 2 * a + 1] = 0;
      sums[l * 3 + a] = 0;
    }
  }
  const int64_t V = s0 * s1 * s2;
  for (int64_t i = 0; i < V; i++) {
    uint64_t l;
    switch (dt) {
      case DT_U8: l = ((const uint8_t*)labels)[i]; break;
      case DT_U16: l = ((const uint16_t*)labels)[i]; break;
      case DT_U32: l = ((const uint32_t*)labels)[i]; break;
      default: l = ((const uint64_t*)labels)[i]; break;
    }
    if (l > N) return 1;
    int64_t c[3];
    if (f_order) {
      c[0] = i % s0;
      c[1] = (i / s0) % s1;
      c[2] = i / (s0 * s1);
    } else {
      c[2] = i % s2;
      c[1] = (i / s2) % s1;
      c[0] = i / (s2 * s1);
    }
    counts[l] += 1u;
    for (int a = 0; a < 3; a++) {
      const uint32_t v = (uint32_t)c[a];
      if (v < bbox[l * 6 + 2 * a]) bbox[l * 6 + 2 * a] = v;
      if (v > bbox[l * 6 + 2 * a + 1]) bbox[l * 6 + 2 * a + 1] = v;
      sums[l * 3 + a] += (uint64_t)c[a];
    }
  }
  return 0;
}

/*
 * oracle_color_vcg: cc3d.color_connectivity_graph restated
 * (cc3d/cc3d_graphs.hpp:525-1106, bound at cc3d/fastcc3d.pyx:941-1018).
 * Every voxel is a node; voxel p is joined to a backward neighbour q when
 * p's own bit for the direction p -> q is set and q lies in the volume (the
 * graph is taken as undirected, so q's bits are not read).  Backward
 * neighbours per variant (the reference's first and second passes):
 *   3D 26 (uint32): (-1,0,0) (-1,-1,0) (0,-1,0) (+1,-1,0) and the 9 voxels
 *                   of plane z - 1;
 *   3D 6:           (-1,0,0) (0,-1,0) (0,0,-1);
 *   2D 8:           (-1,0) (-1,-1) (0,-1) (+1,-1);
 *   2D 4:           (-1,0) (0,-1).
 * Bit of a direction: the voxel_connectivity_graph layout -- for uint32 the
 * 3D one (in 2D the in-plane diagonals are bits 6-9), for uint8 in 2D bits
 * 4-7 for the diagonals.  Components are numbered by their first voxel in
 * raster order (x fastest), the order in which the reference creates and
 * then renumbers its provisional labels.  variant: 0 = 3D 26, 1 = 3D 6,
 * 2 = 2D 8, 3 = 2D 4.  Returns N.
 */
static int vcg_bit3(int dx, int dy, int dz) {
  static const int8_t dirs[26][3] = {
    {+1, 0, 0},   {-1, 0, 0},   {0, +1, 0},   {0, -1, 0},   {0, 0, +1},   {0, 0, -1},   {+1, +1, 0},
    {-1, +1, 0},  {+1, -1, 0},  {-1, -1, 0},  {+1, 0, +1},  {-1, 0, +1},  {0, +1, +1},  {0, -1, +1},
    {+1, 0, -1},  {-1, 0, -1},  {0, +1, -1},  {0, -1, -1},  {+1, +1, +1}, {-1, +1, +1}, {+1, -1, +1},
    {-1, -1, +1}, {+1, +1, -1}, {-1, +1, -1}, {+1, -1, -1}, {-1, -1, -1},
  };
  for (int i = 0; i < 26; i++)
    if (dirs[i][0] == dx && dirs[i][1] == dy && dirs[i][2] == dz) return i;
  return -1;
}
static int vcg_bit2(int dx, int dy) {
  static const int8_t dirs[8][2] = {{+1, 0}, {-1, 0}, {0, +1}, {0, -1}, {+1, +1}, {-1, +1}, {+1, -1}, {-1, -1}};
  for (int i = 0; i < 8; i++)
    if (dirs[i][0] == dx && dirs[i][1] == dy) return i;
  return -1;
}

uint64_t oracle_color_vcg(const void* vcg, int vcg_size, int64_t sx, int64_t sy, int64_t sz, int variant,
                          uint32_t* out) {
  const int64_t V = sx * sy * sz;
  int nb = 0;
  int off[13][3], bit[13];
  if (variant == 0) {
    const int in_plane[4][2] = {{-1, 0}, {-1, -1}, {0, -1}, {+1, -1}};
    for (int i = 0; i < 4; i++) {
      off[nb][0] = in_plane[i][0]; off[nb][1] = in_plane[i][1]; off[nb][2] = 0;
      bit[nb] = vcg_bit3(in_plane[i][0], in_plane[i][1], 0);
      nb++;
    }
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) {
        off[nb][0] = dx; off[nb][1] = dy; off[nb][2] = -1;
        bit[nb] = vcg_bit3(dx, dy, -1);
        nb++;
      }
  } else if (variant == 1) {
    const int o[3][3] = {{-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};
    for (int i = 0; i < 3; i++) {
      off[nb][0] = o[i][0]; off[nb][1] = o[i][1]; off[nb][2] = o[i][2];
      bit[nb] = vcg_bit3(o[i][0], o[i][1], o[i][2]);
      nb++;
    }
  } else {
    const int o8[4][2] = {{-1, 0}, {-1, -1}, {0, -1}, {+1, -1}};
    const int n = variant == 2 ? 4 : 2;
    const int o4[2][2] = {{-1, 0}, {0, -1}};
    for (int i = 0; i < n; i++) {
      const int dx = variant == 2 ? o8[i][0] : o4[i][0], dy = variant == 2 ? o8[i][1] : o4[i][1];
      off[nb][0] = dx; off[nb][1] = dy; off[nb][2] = 0;
      bit[nb] = vcg_size == 1 ? vcg_bit2(dx, dy) : vcg_bit3(dx, dy, 0);
      nb++;
    }
  }
  int64_t* par = (int64_t*)malloc((size_t)V * sizeof(int64_t));
  for (int64_t i = 0; i < V; i++) par[i] = i;
  for (int64_t z = 0; z < sz; z++)
    for (int64_t y = 0; y < sy; y++)
      for (int64_t x = 0; x < sx; x++) {
        const int64_t p = x + sx * (y + sy * z);
        const uint32_t bits = vcg_size == 1 ? ((const uint8_t*)vcg)[p] : ((const uint32_t*)vcg)[p];
        for (int k = 0; k < nb; k++) {
          const int64_t qx = x + off[k][0], qy = y + off[k][1], qz = z + off[k][2];
          if (qx < 0 || qx >= sx || qy < 0 || qy >= sy || qz < 0 || qz >= sz) continue;
          if (!((bits >> bit[k]) & 1u)) continue;
          int64_t a = p, b = qx + sx * (qy + sy * qz);
          while (par[a] != a) a = par[a] = par[par[a]];
          while (par[b] != b) b = par[b] = par[par[b]];
          if (a == b) continue;
          if (a < b) par[b] = a;
          else par[a] = b;
        }
      }
  uint64_t N = 0;
  for (int64_t i = 0; i < V; i++) {
    int64_t r = i;
    while (par[r] != r) r = par[r];
    if (r == i) out[i] = (uint32_t)(++N);
    else out[i] = out[r];
  }
  free(par);
  return N;
}

/*
 * oracle_contacts: cc3d.contacts restated (cc3d/cc3d_graphs.hpp:238-468,
 * bound at cc3d/fastcc3d.pyx:1198-1252).  The volume is walked in raster
 * order (x fastest); for each nonzero voxel the first connectivity / 2
 * backward offsets are visited, formed as the reference forms them:
 *   3D: -x, -y, -z; the in-plane and z diagonals (only when both of their
 *       directions exist); the four z corners, tested on y and z only, so on
 *       the x borders they fall back to the (0, -1, -1) / (0, +1, -1) edge;
 *   2D: -x, -y, (-x) + (-y), (+x) + (-y), each direction present or 0, with
 *       no test of both -- so on the borders a diagonal becomes a face
 *       neighbour again, or on row 0 the forward neighbour (x + 1).
 * A neighbour holding another nonzero label adds the offset's area to the
 * float32 sum of the pair (smaller label first; signed types compare
 * signed): 3D faces wy*wz, wx*wz, wx*wy, 2D faces wy, wx, other offsets 0;
 * with surface_area 0 every offset adds 1.  A pair first seen through an
 * offset of area 0 still exists with sum 0.  lab: labels widened to 64 bits
 * (sign-extended when is_signed).  Returns the number of pairs (written to
 * lo/hi/val, at most cap), or -1 when cap is too small.
 */
typedef struct {
  uint64_t lo, hi;
  float v;
  int used;
} PairCell;

static int pair_less(uint64_t a, uint64_t b, int is_signed) {
  return is_signed ? (int64_t)a < (int64_t)b : a < b;
}

int64_t oracle_contacts(const uint64_t* lab, int is_signed, int64_t sx, int64_t sy, int64_t sz, int connectivity,
                        int surface_area, float wx, float wy, float wz, uint64_t* lo, uint64_t* hi, float* val,
                        int64_t cap) {
  const int d2 = connectivity == 4 || connectivity == 8;
  float areas[13];
  if (d2) {
    areas[0] = wy; areas[1] = wx; areas[2] = 0.0f; areas[3] = 0.0f;
  } else {
    for (int i = 3; i < 13; i++) areas[i] = 0.0f;
    areas[0] = wy * wz; areas[1] = wx * wz; areas[2] = wx * wy;
  }
  if (!surface_area)
    for (int i = 0; i < 13; i++) areas[i] = 1.0f;
  int64_t hcap = 1024;
  while (hcap < 4 * cap) hcap <<= 1;
  PairCell* h = (PairCell*)calloc((size_t)hcap, sizeof(PairCell));
  int64_t n = 0;
  const int64_t sxy = sx * sy;
  const int nn = connectivity / 2;
  for (int64_t z = 0; z < sz; z++)
    for (int64_t y = 0; y < sy; y++)
      for (int64_t x = 0; x < sx; x++) {
        const int64_t loc = x + sx * y + sxy * z;
        const uint64_t cur = lab[loc];
        if (cur == 0) continue;
        int64_t nb[13];
        const int64_t px = x < sx - 1, mx = -(int64_t)(x > 0);
        const int64_t py = y < sy - 1 ? sx : 0, my = y > 0 ? -sx : 0, mz = z > 0 ? -sxy : 0;
        if (d2) {
          nb[0] = mx; nb[1] = my; nb[2] = mx + my; nb[3] = px + my;
        } else {
          nb[0] = mx; nb[1] = my; nb[2] = mz;
          nb[3] = (mx && my) ? mx + my : 0;
          nb[4] = (px && my) ? px + my : 0;
          nb[5] = (mx && mz) ? mx + mz : 0;
          nb[6] = (px && mz) ? px + mz : 0;
          nb[7] = (my && mz) ? my + mz : 0;
          nb[8] = (py && mz) ? py + mz : 0;
          nb[9] = (my && mz) ? mx + my + mz : 0;
          nb[10] = (my && mz) ? px + my + mz : 0;
          nb[11] = (py && mz) ? mx + py + mz : 0;
          nb[12] = (py && mz) ? px + py + mz : 0;
        }
        for (int i = 0; i < nn; i++) {
          const uint64_t v = lab[loc + nb[i]];
          if (v == 0 || v == cur) continue;
          const uint64_t a = pair_less(cur, v, is_signed) ? cur : v, b = pair_less(cur, v, is_signed) ? v : cur;
          uint64_t k = (a * 0x9E3779B97F4A7C15ull) ^ (b + 0x7F4A7C15ull + (a << 6) + (a >> 2));
          int64_t s = (int64_t)(k & (uint64_t)(hcap - 1));
          while (h[s].used && !(h[s].lo == a && h[s].hi == b)) s = (s + 1) & (hcap - 1);
          if (!h[s].used) {
            if (n >= cap) {
              free(h);
              return -1;
            }
            h[s].used = 1;
            h[s].lo = a;
            h[s].hi = b;
            h[s].v = 0.0f;
            n++;
          }
          h[s].v += areas[i];
        }
      }
  int64_t m = 0;
  for (int64_t s = 0; s < hcap; s++)
    if (h[s].used) {
      lo[m] = h[s].lo;
      hi[m] = h[s].hi;
      val[m] = h[s].v;
      m++;
    }
  free(h);
  return m;
}

/*
 * oracle_runs: cc3d's runs() restated (cc3d/cc3d_graphs.hpp:470-501): the
 * maximal runs of equal values of the flat array, as (label, start, end),
 * nonzero labels only -- except that a one-voxel array gives its single run
 * whatever its value.  Runs come out in position order.  Returns the count.
 */
int64_t oracle_runs(const uint64_t* lab, int64_t n, uint64_t* label, uint64_t* start, uint64_t* end) {
  if (n == 0) return 0;
  if (n == 1) {
    label[0] = lab[0]; start[0] = 0; end[0] = 1;
    return 1;
  }
  int64_t m = 0, s = 0;
  uint64_t cur = lab[0];
  for (int64_t i = 1; i < n; i++) {
    if (lab[i] != cur) {
      if (cur != 0) {
        label[m] = cur; start[m] = (uint64_t)s; end[m] = (uint64_t)i; m++;
      }
      cur = lab[i];
      s = i;
    }
  }
  if (cur != 0) {
    label[m] = cur; start[m] = (uint64_t)s; end[m] = (uint64_t)n; m++;
  }
  return m;
}
