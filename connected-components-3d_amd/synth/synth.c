/*
 * synth.c -- deterministic synthetic volumes for the labelling benchmarks.
 *
 * Every random draw is a counter-based hash (SplitMix64 finaliser, Steele,
 * Lea & Flood 2014) of (case seed, stream, index), so the bytes depend only on
 * the arguments -- not on thread count, machine or library versions.
 *
 * Volumes are x fastest: index = x + sx * (y + sy * z).
 *
 *   voronoi:  seeds uniform on integer voxel coordinates; a voxel belongs to
 *             the seed of minimum integer squared distance, ties to the lowest
 *             seed index.  Membranes: a voxel with a different-region 26-
 *             neighbour (L-inf 1, in bounds) is set to 0 with probability
 *             p_membrane.  Salt: each voxel is then overwritten, with
 *             probability p_salt, by the ID of a uniformly drawn seed.
 *   bernoulli: foreground (1) with probability p.
 *   tones:    Voronoi regions set to a tone drawn from a list, plus i.i.d.
 *             uniform float32 noise in [-a, a].
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static inline uint64_t draw(uint64_t seed, uint64_t stream, uint64_t i) {
  return mix64(mix64(seed ^ (stream * 0xD1B54A32D192ED03ull)) + i);
}
/* true with probability p (53-bit fixed point comparison) */
static inline int chance(uint64_t r, double p) {
  const uint64_t thr = (uint64_t)(p * 9007199254740992.0);
  return (r >> 11) < thr;
}

enum { S_SEEDX = 1, S_SEEDY, S_SEEDZ, S_ID, S_MEMB, S_SALT, S_SALTID, S_BERN, S_TONE, S_NOISE };

typedef struct {
  int64_t sx, sy, sz, nseeds;
  const int32_t* seeds; /* 3 * nseeds */
  int64_t cell;
  int64_t cx, cy, cz;
  const int64_t* cstart; /* cells + 1 */
  const int32_t* cidx;
  uint32_t* region; /* output: seed index per voxel */
  int64_t z0, z1;
} vor_job;

static void* vor_worker(void* arg) {
  vor_job* j = (vor_job*)arg;
  const int64_t sx = j->sx, sy = j->sy, c = j->cell;
  for (int64_t z = j->z0; z < j->z1; z++)
    for (int64_t y = 0; y < sy; y++)
      for (int64_t x = 0; x < sx; x++) {
        const int64_t gx = x / c, gy = y / c, gz = z / c;
        int64_t best = INT64_MAX;
        int64_t bi = -1;
        for (int64_t r = 0;; r++) {
          if (r >= 1) {
            const int64_t lb = (r - 1) * c;
            if (best < lb * lb) break;
          }
          if (r > j->cx + j->cy + j->cz) break;
          for (int64_t dz = -r; dz <= r; dz++) {
            const int64_t qz = gz + dz;
            if (qz < 0 || qz >= j->cz) continue;
            for (int64_t dy = -r; dy <= r; dy++) {
              const int64_t qy = gy + dy;
              if (qy < 0 || qy >= j->cy) continue;
              const int on_shell_yz = (llabs(dz) == r || llabs(dy) == r);
              for (int64_t dx = -r; dx <= r; dx++) {
                if (!on_shell_yz && llabs(dx) != r) continue;
                const int64_t qx = gx + dx;
                if (qx < 0 || qx >= j->cx) continue;
                const int64_t cell = qx + j->cx * (qy + j->cy * qz);
                for (int64_t k = j->cstart[cell]; k < j->cstart[cell + 1]; k++) {
                  const int32_t s = j->cidx[k];
                  const int64_t ddx = x - j->seeds[3 * s], ddy = y - j->seeds[3 * s + 1],
                                ddz = z - j->seeds[3 * s + 2];
                  const int64_t d = ddx * ddx + ddy * ddy + ddz * ddz;
                  if (d < best || (d == best && s < bi)) {
                    best = d;
                    bi = s;
                  }
                }
              }
            }
          }
        }
        j->region[x + sx * (y + sy * z)] = (uint32_t)bi;
      }
  return NULL;
}

/* Seed coordinates for (seed, nseeds). */
static void make_seeds(uint64_t seed, int64_t sx, int64_t sy, int64_t sz, int64_t n, int32_t* s) {
  for (int64_t i = 0; i < n; i++) {
    s[3 * i] = (int32_t)(draw(seed, S_SEEDX, i) % (uint64_t)sx);
    s[3 * i + 1] = (int32_t)(draw(seed, S_SEEDY, i) % (uint64_t)sy);
    s[3 * i + 2] = (int32_t)(draw(seed, S_SEEDZ, i) % (uint64_t)sz);
  }
}

/* Voronoi region index per voxel (threads > 0 splits z-slabs). Returns 0. */
int synth_voronoi_regions(uint64_t seed, int64_t sx, int64_t sy, int64_t sz, int64_t nseeds,
                          uint32_t* region, int threads) {
  if (nseeds <= 0 || sx <= 0 || sy <= 0 || sz <= 0) return -1;
  int32_t* seeds = (int32_t*)malloc(sizeof(int32_t) * 3 * nseeds);
  make_seeds(seed, sx, sy, sz, nseeds, seeds);
  const double vol = (double)sx * sy * sz;
  int64_t cell = (int64_t)ceil(cbrt(vol / (double)nseeds * 2.0));
  if (cell < 1) cell = 1;
  const int64_t cx = (sx + cell - 1) / cell, cy = (sy + cell - 1) / cell, cz = (sz + cell - 1) / cell;
  const int64_t ncell = cx * cy * cz;
  int64_t* cstart = (int64_t*)calloc(ncell + 1, sizeof(int64_t));
  int32_t* cidx = (int32_t*)malloc(sizeof(int32_t) * nseeds);
  for (int64_t i = 0; i < nseeds; i++) {
    const int64_t c = seeds[3 * i] / cell + cx * (seeds[3 * i + 1] / cell + cy * (seeds[3 * i + 2] / cell));
    cstart[c + 1]++;
  }
  for (int64_t c = 0; c < ncell; c++) cstart[c + 1] += cstart[c];
  int64_t* fill = (int64_t*)calloc(ncell, sizeof(int64_t));
  for (int64_t i = 0; i < nseeds; i++) {
    const int64_t c = seeds[3 * i] / cell + cx * (seeds[3 * i + 1] / cell + cy * (seeds[3 * i + 2] / cell));
    cidx[cstart[c] + fill[c]++] = (int32_t)i;
  }
  if (threads < 1) threads = 1;
  if (threads > 64) threads = 64;
  pthread_t th[64];
  vor_job jobs[64];
  for (int t = 0; t < threads; t++) {
    vor_job* j = &jobs[t];
    j->sx = sx; j->sy = sy; j->sz = sz; j->nseeds = nseeds; j->seeds = seeds;
    j->cell = cell; j->cx = cx; j->cy = cy; j->cz = cz; j->cstart = cstart; j->cidx = cidx;
    j->region = region;
    j->z0 = sz * t / threads;
    j->z1 = sz * (t + 1) / threads;
    pthread_create(&th[t], NULL, vor_worker, j);
  }
  for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
  free(seeds); free(cstart); free(cidx); free(fill);
  return 0;
}

/* Seed IDs: uniform in [lo, hi) (hi = 0 means the full 64-bit range, never 0). */
static uint64_t seed_id(uint64_t seed, int64_t i, uint64_t lo, uint64_t hi) {
  const uint64_t r = draw(seed, S_ID, (uint64_t)i);
  if (hi == 0) return r ? r : 1;
  return lo + r % (hi - lo);
}

/* Voronoi label volume with membranes and salt; out is uint32 (width 4) or
 * uint64 (width 8). */
int synth_voronoi(uint64_t seed, int64_t sx, int64_t sy, int64_t sz, int64_t nseeds, uint64_t id_lo,
                  uint64_t id_hi, double p_membrane, double p_salt, void* out, int width, int threads) {
  const int64_t V = sx * sy * sz;
  uint32_t* region = (uint32_t*)malloc(sizeof(uint32_t) * V);
  if (!region) return -1;
  if (synth_voronoi_regions(seed, sx, sy, sz, nseeds, region, threads)) {
    free(region);
    return -1;
  }
  uint64_t* ids = (uint64_t*)malloc(sizeof(uint64_t) * nseeds);
  for (int64_t i = 0; i < nseeds; i++) ids[i] = seed_id(seed, i, id_lo, id_hi);
  for (int64_t z = 0; z < sz; z++)
    for (int64_t y = 0; y < sy; y++)
      for (int64_t x = 0; x < sx; x++) {
        const int64_t i = x + sx * (y + sy * z);
        const uint32_t r = region[i];
        uint64_t v = ids[r];
        int boundary = 0;
        for (int dz = -1; dz <= 1 && !boundary; dz++)
          for (int dy = -1; dy <= 1 && !boundary; dy++)
            for (int dx = -1; dx <= 1; dx++) {
              const int64_t qx = x + dx, qy = y + dy, qz = z + dz;
              if (qx < 0 || qy < 0 || qz < 0 || qx >= sx || qy >= sy || qz >= sz) continue;
              if (region[qx + sx * (qy + sy * qz)] != r) {
                boundary = 1;
                break;
              }
            }
        if (boundary && chance(draw(seed, S_MEMB, (uint64_t)i), p_membrane)) v = 0;
        if (chance(draw(seed, S_SALT, (uint64_t)i), p_salt))
          v = ids[draw(seed, S_SALTID, (uint64_t)i) % (uint64_t)nseeds];
        if (width == 4) ((uint32_t*)out)[i] = (uint32_t)v;
        else ((uint64_t*)out)[i] = v;
      }
  free(ids);
  free(region);
  return 0;
}

/* Bernoulli(p) uint8 volume. */
int synth_bernoulli(uint64_t seed, int64_t V, double p, uint8_t* out) {
  for (int64_t i = 0; i < V; i++) out[i] = (uint8_t)chance(draw(seed, S_BERN, (uint64_t)i), p);
  return 0;
}

/* Voronoi regions with a tone per seed drawn from tones[ntones], plus float32
 * noise in [-amp, amp]. */
int synth_tones(uint64_t seed, int64_t sx, int64_t sy, int64_t sz, int64_t nseeds, const float* tones,
                int ntones, float amp, float* out, int threads) {
  const int64_t V = sx * sy * sz;
  uint32_t* region = (uint32_t*)malloc(sizeof(uint32_t) * V);
  if (!region) return -1;
  if (synth_voronoi_regions(seed, sx, sy, sz, nseeds, region, threads)) {
    free(region);
    return -1;
  }
  for (int64_t i = 0; i < V; i++) {
    const float tone = tones[draw(seed, S_TONE, region[i]) % (uint64_t)ntones];
    const float u = (float)(draw(seed, S_NOISE, (uint64_t)i) >> 40) * (1.0f / 16777216.0f);
    const float noise = (2.0f * amp) * u - amp;
    out[i] = tone + noise;
  }
  free(region);
  return 0;
}
