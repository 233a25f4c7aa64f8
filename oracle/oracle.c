/*
 * oracle.c -- CPU restatement of nvpro-pipeline's transform propagation and
 * frustum culling.  TEST INFRASTRUCTURE ONLY (see oracle.h): the product
 * library never links or calls it.
 *
 * Must be compiled with -ffp-contract=off and without -ffast-math: each
 * statement below is one IEEE-754 binary32 operation, in the reference's
 * evaluation order.
 */
#include "oracle.h"

#include <float.h>
#include <pthread.h>
#include <stdint.h>
#include <string.h>

/* dp/math/Matmnt.h:1371-1379 */
void orc_vec4_mul_mat44(const float v[4], const float m[16], float out[4])
{
  float r[4];
  for (int c = 0; c < 4; ++c) {
    float s = v[0] * m[0 * 4 + c];
    s = s + v[1] * m[1 * 4 + c];
    s = s + v[2] * m[2 * 4 + c];
    s = s + v[3] * m[3 * 4 + c];
    r[c] = s;
  }
  memcpy(out, r, sizeof r);
}

/* dp/math/Matmnt.h:1381-1417 */
void orc_mat44_mul(const float a[16], const float b[16], float out[16])
{
  float r[16];
  for (int i = 0; i < 4; ++i) {
    for (int j = 0; j < 4; ++j) {
      float s = a[i * 4 + 0] * b[0 * 4 + j];
      s = s + a[i * 4 + 1] * b[1 * 4 + j];
      s = s + a[i * 4 + 2] * b[2 * 4 + j];
      s = s + a[i * 4 + 3] * b[3 * 4 + j];
      r[i * 4 + j] = s;
    }
  }
  memcpy(out, r, sizeof r);
}

void orc_mat44_mul_f64(const double a[16], const double b[16], double out[16])
{
  double r[16];
  for (int i = 0; i < 4; ++i) {
    for (int j = 0; j < 4; ++j) {
      double s = a[i * 4 + 0] * b[0 * 4 + j];
      s = s + a[i * 4 + 1] * b[1 * 4 + j];
      s = s + a[i * 4 + 2] * b[2 * 4 + j];
      s = s + a[i * 4 + 3] * b[3 * 4 + j];
      r[i * 4 + j] = s;
    }
  }
  memcpy(out, r, sizeof r);
}

/* dp/culling/cpu/src/ManagerImpl.cpp:131 (point) and :152-156 (scalar branch). */
void orc_obb_build(const float lower4[4], const float extent4[4], const float m[16], float obb[16])
{
  orc_vec4_mul_mat44(lower4, m, obb);
  for (int c = 0; c < 4; ++c) {
    obb[4 + c] = extent4[0] * m[0 * 4 + c];
    obb[8 + c] = extent4[1] * m[1 * 4 + c];
    obb[12 + c] = extent4[2] * m[2 * 4 + c];
  }
}

static void vadd4(const float* a, const float* b, float* out)
{
  for (int c = 0; c < 4; ++c) out[c] = a[c] + b[c];
}

/* dp/culling/cpu/src/ManagerImpl.cpp:268-281 */
void orc_obb_corners_clip(const float vp[16], const float obb[16], float v[32])
{
  float x[4], y[4], z[4];
  orc_vec4_mul_mat44(obb + 0, vp, v + 0);
  orc_vec4_mul_mat44(obb + 4, vp, x);
  orc_vec4_mul_mat44(obb + 8, vp, y);
  orc_vec4_mul_mat44(obb + 12, vp, z);
  vadd4(v + 0, x, v + 4);   /* vectors[1] = vectors[0] + x */
  vadd4(v + 0, y, v + 8);   /* vectors[2] = vectors[0] + y */
  vadd4(v + 4, y, v + 12);  /* vectors[3] = vectors[1] + y */
  vadd4(v + 0, z, v + 16);  /* vectors[4] = vectors[0] + z */
  vadd4(v + 4, z, v + 20);  /* vectors[5] = vectors[1] + z */
  vadd4(v + 8, z, v + 24);  /* vectors[6] = vectors[2] + z */
  vadd4(v + 12, z, v + 28); /* vectors[7] = vectors[3] + z */
}

/* dp/culling/cpu/src/ManagerImpl.cpp:199-229 */
static void cull_flags(const float* p, unsigned* cfo, unsigned* cfa)
{
  unsigned cf = 0;
  if (p[0] <= -p[3]) cf |= 0x01u;
  else if (p[3] <= p[0]) cf |= 0x02u;
  if (p[1] <= -p[3]) cf |= 0x04u;
  else if (p[3] <= p[1]) cf |= 0x08u;
  if (p[2] <= -p[3]) cf |= 0x10u;
  else if (p[3] <= p[2]) cf |= 0x20u;
  *cfo |= cf;
  *cfa &= cf;
}

/* dp/culling/cpu/src/ManagerImpl.cpp:263-289 */
int orc_is_visible_obb(const float vp[16], const float obb[16])
{
  float v[32];
  unsigned cfo = 0, cfa = ~0u;
  orc_obb_corners_clip(vp, obb, v);
  for (int k = 0; k < 8; ++k) cull_flags(v + 4 * k, &cfo, &cfa);
  return (!cfo || !cfa) ? 1 : 0;
}

static void set_bit(uint32_t* mask, size_t i, int value)
{
  /* 64-bit BitArray words viewed as little-endian uint32 words: bit i lives in
   * uint32 word i >> 5 at position i & 31. */
  if (value) mask[i >> 5] |= (uint32_t)1u << (i & 31);
  else mask[i >> 5] &= ~((uint32_t)1u << (i & 31));
}

static size_t mask_words(size_t n) { return 2 * ((n + 63) / 64); }

void orc_cull_obbs(size_t n, const float* obbs, const float vp[16], uint32_t* mask)
{
  /* dp::util::BitArray visible(count) starts cleared (BitArray.cpp:45-50). */
  memset(mask, 0, mask_words(n) * sizeof(uint32_t));
  for (size_t i = 0; i < n; ++i) set_bit(mask, i, orc_is_visible_obb(vp, obbs + 16 * i));
}

static const float* matrix_at(const void* base, size_t stride, uint64_t index)
{
  return (const float*)((const char*)base + index * stride);
}

/* dp/culling/cpu/src/ManagerImpl.cpp:114-162 */
void orc_obb_build_group(size_t n, const float* lower4, const float* extent4, const uint64_t* xform,
                         const void* matrices, size_t stride, float* obbs)
{
  for (size_t i = 0; i < n; ++i)
    orc_obb_build(lower4 + 4 * i, extent4 + 4 * i, matrix_at(matrices, stride, xform[i]), obbs + 16 * i);
}

/* dp/culling/cpu/src/ManagerImpl.cpp:467-518 with GroupCPU::updateOBBs (:114-162). */
void orc_cull_group(size_t n, const float* lower4, const float* extent4, const uint64_t* xform,
                    const void* matrices, size_t stride, const float vp[16], uint32_t* mask)
{
  memset(mask, 0, mask_words(n) * sizeof(uint32_t));
  for (size_t i = 0; i < n; ++i) {
    float obb[16];
    orc_obb_build(lower4 + 4 * i, extent4 + 4 * i, matrix_at(matrices, stride, xform[i]), obb);
    set_bit(mask, i, orc_is_visible_obb(vp, obb));
  }
}

/* dp/math/Boxnt.h:245-258, Box4f::update */
static void box4_update(float* lo, float* hi, const float* p)
{
  for (int i = 0; i < 4; ++i) {
    if (lo[i] > p[i]) lo[i] = p[i];
    if (hi[i] < p[i]) hi[i] = p[i];
  }
}

/* dp/culling/src/ManagerBitSet.cpp:268-306 (scalar branch) */
void orc_bbox_group(size_t n, const float* lower4, const float* extent4, const uint64_t* xform,
                    const void* matrices, size_t stride, float out6[6])
{
  float lo[4], hi[4];
  for (int i = 0; i < 4; ++i) { lo[i] = FLT_MAX; hi[i] = -FLT_MAX; }  /* Boxnt::init, Boxnt.h:226-233 */
  for (size_t k = 0; k < n; ++k) {
    const float* m = matrix_at(matrices, stride, xform[k]);
    const float* e = extent4 + 4 * k;
    float v[32], x[4], y[4], z[4];
    orc_vec4_mul_mat44(lower4 + 4 * k, m, v);
    for (int c = 0; c < 4; ++c) {
      x[c] = e[0] * m[0 * 4 + c];
      y[c] = e[1] * m[1 * 4 + c];
      z[c] = e[2] * m[2 * 4 + c];
    }
    vadd4(v + 0, x, v + 4);
    vadd4(v + 0, y, v + 8);
    vadd4(v + 4, y, v + 12);
    vadd4(v + 0, z, v + 16);
    vadd4(v + 4, z, v + 20);
    vadd4(v + 8, z, v + 24);
    vadd4(v + 12, z, v + 28);
    for (int c = 0; c < 8; ++c) box4_update(lo, hi, v + 4 * c);
  }
  /* return Box3f(lower.xyz, upper.xyz): Boxnt(p0, p1) = init(); update(p0); update(p1) (Boxnt.h:213-219). */
  float blo[4] = {FLT_MAX, FLT_MAX, FLT_MAX, 0}, bhi[4] = {-FLT_MAX, -FLT_MAX, -FLT_MAX, 0};
  float p0[4] = {lo[0], lo[1], lo[2], 0}, p1[4] = {hi[0], hi[1], hi[2], 0};
  for (int i = 0; i < 3; ++i) {
    if (blo[i] > p0[i]) blo[i] = p0[i];
    if (bhi[i] < p0[i]) bhi[i] = p0[i];
  }
  for (int i = 0; i < 3; ++i) {
    if (blo[i] > p1[i]) blo[i] = p1[i];
    if (bhi[i] < p1[i]) bhi[i] = p1[i];
  }
  for (int i = 0; i < 3; ++i) { out6[i] = blo[i]; out6[3 + i] = bhi[i]; }
}

/* dp/transform/src/Tree.cpp:133-166 */
void orc_tree_compute(size_t n_levels, const uint64_t* level_offsets, const uint32_t* entries,
                      const float* local, float* world, size_t n_index,
                      uint8_t* dirty_local, uint8_t* dirty_world, uint8_t* dirty_world_event)
{
  for (size_t l = 0; l < n_levels; ++l) {
    for (uint64_t e = level_offsets[l]; e < level_offsets[l + 1]; ++e) {
      uint32_t parent = entries[2 * e + 0];
      uint32_t t = entries[2 * e + 1];
      if (dirty_world[parent] || dirty_local[t]) {
        orc_mat44_mul(local + 16 * (size_t)t, world + 16 * (size_t)parent, world + 16 * (size_t)t);
        dirty_world[t] = 1;
      }
    }
  }
  if (dirty_world_event) memcpy(dirty_world_event, dirty_world, n_index);
  memset(dirty_local, 0, n_index);
  memset(dirty_world, 0, n_index);
}

void orc_tree_compute_f64(size_t n_levels, const uint64_t* level_offsets, const uint32_t* entries,
                          const double* local, double* world)
{
  for (size_t l = 0; l < n_levels; ++l) {
    for (uint64_t e = level_offsets[l]; e < level_offsets[l + 1]; ++e) {
      uint32_t parent = entries[2 * e + 0];
      uint32_t t = entries[2 * e + 1];
      orc_mat44_mul_f64(local + 16 * (size_t)t, world + 16 * (size_t)parent, world + 16 * (size_t)t);
    }
  }
}

/* dp/culling/src/ResultBitSet.cpp:100-107, dp/util/BitArray.h:104-137 */
size_t orc_result_diff(size_t n, const uint32_t* old_mask, const uint32_t* new_mask, uint32_t* changed_idx)
{
  size_t count = 0;
  for (size_t i = 0; i < n; ++i) {
    uint32_t a = (old_mask[i >> 5] >> (i & 31)) & 1u;
    uint32_t b = (new_mask[i >> 5] >> (i & 31)) & 1u;
    if (a ^ b) changed_idx[count++] = (uint32_t)i;
  }
  return count;
}

/* dp/math/Matmnt.h:1171-1201.  m[4*r + c] = M[r][c].  Every expression is
 * evaluated left to right as in the reference: a - b + c = (a - b) + c, and
 * -a * b = (-a) * b (the sign flip is exact). */
int orc_invert_transpose(const float m[16], float out[16])
{
#define M(r, c) m[4 * (r) + (c)]
  const float s0 = M(0, 0) * M(1, 1) - M(0, 1) * M(1, 0);
  const float c5 = M(2, 2) * M(3, 3) - M(2, 3) * M(3, 2);
  const float s1 = M(0, 0) * M(1, 2) - M(0, 2) * M(1, 0);
  const float c4 = M(2, 1) * M(3, 3) - M(2, 3) * M(3, 1);
  const float s2 = M(0, 0) * M(1, 3) - M(0, 3) * M(1, 0);
  const float c3 = M(2, 1) * M(3, 2) - M(2, 2) * M(3, 1);
  const float s3 = M(0, 1) * M(1, 2) - M(0, 2) * M(1, 1);
  const float c2 = M(2, 0) * M(3, 3) - M(2, 3) * M(3, 0);
  const float s4 = M(0, 1) * M(1, 3) - M(0, 3) * M(1, 1);
  const float c1 = M(2, 0) * M(3, 2) - M(2, 2) * M(3, 0);
  const float s5 = M(0, 2) * M(1, 3) - M(0, 3) * M(1, 2);
  const float c0 = M(2, 0) * M(3, 1) - M(2, 1) * M(3, 0);
  float det = s0 * c5;
  det = det - s1 * c4;
  det = det + s2 * c3;
  det = det + s3 * c2;
  det = det - s4 * c1;
  det = det + s5 * c0;
  if (!(det != 0.0f)) return 0;
  const float inv = 1.0f / det;
  float r[16];
  /* r[4*i + j] = mOut[i][j] */
  r[0 * 4 + 0] = ((M(1, 1) * c5 - M(1, 2) * c4) + M(1, 3) * c3) * inv;
  r[1 * 4 + 0] = ((-M(0, 1) * c5 + M(0, 2) * c4) - M(0, 3) * c3) * inv;
  r[2 * 4 + 0] = ((M(3, 1) * s5 - M(3, 2) * s4) + M(3, 3) * s3) * inv;
  r[3 * 4 + 0] = ((-M(2, 1) * s5 + M(2, 2) * s4) - M(2, 3) * s3) * inv;
  r[0 * 4 + 1] = ((-M(1, 0) * c5 + M(1, 2) * c2) - M(1, 3) * c1) * inv;
  r[1 * 4 + 1] = ((M(0, 0) * c5 - M(0, 2) * c2) + M(0, 3) * c1) * inv;
  r[2 * 4 + 1] = ((-M(3, 0) * s5 + M(3, 2) * s2) - M(3, 3) * s1) * inv;
  r[3 * 4 + 1] = ((M(2, 0) * s5 - M(2, 2) * s2) + M(2, 3) * s1) * inv;
  r[0 * 4 + 2] = ((M(1, 0) * c4 - M(1, 1) * c2) + M(1, 3) * c0) * inv;
  r[1 * 4 + 2] = ((-M(0, 0) * c4 + M(0, 1) * c2) - M(0, 3) * c0) * inv;
  r[2 * 4 + 2] = ((M(3, 0) * s4 - M(3, 1) * s2) + M(3, 3) * s0) * inv;
  r[3 * 4 + 2] = ((-M(2, 0) * s4 + M(2, 1) * s2) - M(2, 3) * s0) * inv;
  r[0 * 4 + 3] = ((-M(1, 0) * c3 + M(1, 1) * c1) - M(1, 2) * c0) * inv;
  r[1 * 4 + 3] = ((M(0, 0) * c3 - M(0, 1) * c1) + M(0, 2) * c0) * inv;
  r[2 * 4 + 3] = ((-M(3, 0) * s3 + M(3, 1) * s1) - M(3, 2) * s0) * inv;
  r[3 * 4 + 3] = ((M(2, 0) * s3 - M(2, 1) * s1) + M(2, 2) * s0) * inv;
#undef M
  memcpy(out, r, sizeof r);
  return 1;
}

/* ---- a static partition over host threads (CPU timing only) -----------------
 * The same per-object functions as orc_obb_build_group and orc_cull_obbs, with
 * the object range split into contiguous blocks of whole 64-bit mask words, one
 * per thread.  Used to time the CPU path on every host core; results are
 * identical to the single-threaded functions. */
typedef struct {
  size_t begin, end;
  const float *lower4, *extent4;
  const uint64_t* xform;
  const void* matrices;
  size_t stride;
  const float* vp;
  float* obbs;
  uint32_t* mask;
} orc_range;

static void* obb_range(void* arg)
{
  const orc_range* r = (const orc_range*)arg;
  for (size_t i = r->begin; i < r->end; ++i)
    orc_obb_build(r->lower4 + 4 * i, r->extent4 + 4 * i, matrix_at(r->matrices, r->stride, r->xform[i]),
                  r->obbs + 16 * i);
  return NULL;
}

static void* cull_range(void* arg)
{
  const orc_range* r = (const orc_range*)arg;
  for (size_t i = r->begin; i < r->end; ++i) set_bit(r->mask, i, orc_is_visible_obb(r->vp, r->obbs + 16 * i));
  return NULL;
}

/* The host threads: created on first use and kept for the life of the process,
 * so a call costs a wake-up and a join, not thread creation (the CPU timing
 * then measures the cores).  Worker k (1 <= k < threads) runs part k of each
 * job; the calling thread runs part 0. */
enum { kMaxThreads = 256 };
static struct {
  pthread_mutex_t mu;
  pthread_cond_t go, done;
  int workers;             /* threads created (ids 1..workers) */
  unsigned long job;       /* bumped for every job */
  int pending;             /* workers still inside the current job */
  void* (*fn)(void*);
  orc_range* parts;
  int n_parts;
  unsigned long first_job[kMaxThreads]; /* the job a worker was created for: it runs that one first */
} g_pool = {PTHREAD_MUTEX_INITIALIZER, PTHREAD_COND_INITIALIZER, PTHREAD_COND_INITIALIZER, 0, 0, 0, NULL, NULL, 0, {0}};

static void* pool_worker(void* arg)
{
  const int id = (int)(intptr_t)arg;
  unsigned long seen = 0;
  pthread_mutex_lock(&g_pool.mu);
  seen = g_pool.first_job[id] - 1;
  for (;;) {
    while (g_pool.job == seen) pthread_cond_wait(&g_pool.go, &g_pool.mu);
    seen = g_pool.job;
    void* (*fn)(void*) = g_pool.fn;
    orc_range* part = id < g_pool.n_parts ? &g_pool.parts[id] : NULL;
    pthread_mutex_unlock(&g_pool.mu);
    if (part) fn(part);
    pthread_mutex_lock(&g_pool.mu);
    if (--g_pool.pending == 0) pthread_cond_signal(&g_pool.done);
  }
  return NULL;
}

static pthread_mutex_t g_run_mu = PTHREAD_MUTEX_INITIALIZER; /* one job at a time */

static void run_ranges(size_t n, int threads, orc_range proto, void* (*fn)(void*))
{
  pthread_mutex_lock(&g_run_mu);
  if (threads < 1) threads = 1;
  if (threads > kMaxThreads) threads = kMaxThreads;
  orc_range part[kMaxThreads];
  const size_t words = (n + 63) / 64;
  const size_t per = (words + (size_t)threads - 1) / (size_t)threads;
  for (int t = 0; t < threads; ++t) {
    part[t] = proto;
    part[t].begin = (size_t)t * per * 64;
    part[t].end = ((size_t)t + 1) * per * 64;
    if (part[t].begin > n) part[t].begin = n;
    if (part[t].end > n) part[t].end = n;
  }
  pthread_mutex_lock(&g_pool.mu);
  while (g_pool.workers < threads - 1) {
    pthread_t tid;
    g_pool.first_job[g_pool.workers + 1] = g_pool.job + 1;  /* the job posted below */
    if (pthread_create(&tid, NULL, pool_worker, (void*)(intptr_t)(g_pool.workers + 1)) != 0) break;
    pthread_detach(tid);
    ++g_pool.workers;
  }
  /* parts no worker exists for (thread creation failed) run here */
  const int served = g_pool.workers + 1 < threads ? g_pool.workers + 1 : threads;
  g_pool.fn = fn;
  g_pool.parts = part;
  g_pool.n_parts = served;
  g_pool.pending = g_pool.workers;
  ++g_pool.job;
  pthread_cond_broadcast(&g_pool.go);
  pthread_mutex_unlock(&g_pool.mu);
  fn(&part[0]);
  for (int t = served; t < threads; ++t) fn(&part[t]);
  pthread_mutex_lock(&g_pool.mu);
  while (g_pool.pending > 0) pthread_cond_wait(&g_pool.done, &g_pool.mu);
  pthread_mutex_unlock(&g_pool.mu);
  pthread_mutex_unlock(&g_run_mu);
}

void orc_obb_build_group_mt(size_t n, const float* lower4, const float* extent4, const uint64_t* xform,
                            const void* matrices, size_t stride, float* obbs, int threads)
{
  orc_range r = {0, 0, lower4, extent4, xform, matrices, stride, NULL, obbs, NULL};
  run_ranges(n, threads, r, obb_range);
}

void orc_cull_obbs_mt(size_t n, const float* obbs, const float vp[16], uint32_t* mask, int threads)
{
  memset(mask, 0, mask_words(n) * sizeof(uint32_t));
  orc_range r = {0, 0, NULL, NULL, NULL, NULL, 0, vp, (float*)obbs, mask};
  run_ranges(n, threads, r, cull_range);
}
