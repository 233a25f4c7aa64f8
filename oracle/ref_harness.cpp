// ref_harness.cpp -- C entry points around the reference's own compiled code.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
// the reference sources where they lie (dp/transform/src/Tree.cpp,
// dp/util/src/BitArray.cpp, dp/util/src/Observer.cpp, dp/math/src/Matmnt.cpp)
// into oracle/_ref/libdpref.so.  Those translation units build with the
// image's g++ alone.  The culling translation units (dp/culling/...) include
// Boost headers that this image does not have, so they are not built.
//
// Nothing here restates an algorithm: every function forwards to the
// reference's classes and operators, so that tests can pin the restatement in
// oracle.c against the reference itself.
#include <dp/math/Matmnt.h>
#include <dp/math/Quatt.h>
#include <dp/transform/Tree.h>
#include <dp/util/BitArray.h>

#include <cstring>
#include <stdexcept>
#include <vector>

using dp::math::Mat44f;
using dp::math::Vec3f;
using dp::math::Vec4f;

namespace {

Mat44f load_mat(const float* m)
{
  Mat44f r;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) r[i][j] = m[i * 4 + j];
  return r;
}

void store_mat(const Mat44f& r, float* m)
{
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) m[i * 4 + j] = r[i][j];
}

// Tree subclass: gives the harness read access to the protected level lists
// and captures EventWorldMatricesChanged through the Subject/Observer API.
class RefTree : public dp::transform::Tree, public dp::util::Observer
{
public:
  RefTree() { attach(this); }
  ~RefTree() { detach(this); }

  void onNotify(dp::util::Event const& event, dp::util::Payload*) override
  {
    auto const& e = static_cast<dp::transform::Tree::EventWorldMatricesChanged const&>(event);
    dp::util::BitArray const& bits = e.getDirtyWorldMatrices();
    m_lastEvent.assign(bits.getSize(), 0);
    for (size_t i = 0; i < bits.getSize(); ++i) m_lastEvent[i] = bits.getBit(i) ? 1 : 0;
  }
  void onDestroyed(dp::util::Subject const&, dp::util::Payload*) override {}

  size_t levelCount() const { return m_transformLevels.size(); }
  size_t levelSize(size_t l) const { return m_transformLevels[l].transformListEntries.size(); }
  void levelEntries(size_t l, uint32_t* out) const
  {
    auto const& v = m_transformLevels[l].transformListEntries;
    for (size_t k = 0; k < v.size(); ++k) {
      out[2 * k] = v[k].parent;
      out[2 * k + 1] = v[k].transform;
    }
  }
  uint32_t nodeLevel(uint32_t i) const { return m_level[i]; }
  bool isFree(uint32_t i) const { return m_freeTransforms.getBit(i); }
  std::vector<uint8_t> m_lastEvent;
};

}  // namespace

extern "C" {

// ---- dp/math operators -----------------------------------------------------
void ref_vec4_mul_mat44(const float* v, const float* m, float* out)
{
  Vec4f a(v[0], v[1], v[2], v[3]);
  Vec4f r = a * load_mat(m);
  for (int i = 0; i < 4; ++i) out[i] = r[i];
}

void ref_mat44_mul(const float* a, const float* b, float* out)
{
  store_mat(load_mat(a) * load_mat(b), out);
}

// dp::math::invertTranspose (Matmnt.h:1171-1201): returns its result; out is
// written only when it returns true (the reference leaves mOut untouched otherwise).
int ref_invert_transpose(const float* m, float* out)
{
  Mat44f r = load_mat(out);
  const bool ok = dp::math::invertTranspose(load_mat(m), r);
  if (ok) store_mat(r, out);
  return ok ? 1 : 0;
}

void ref_scalar_mul_vec4(float s, const float* v, float* out)
{
  Vec4f a(v[0], v[1], v[2], v[3]);
  Vec4f r = s * a;
  for (int i = 0; i < 4; ++i) out[i] = r[i];
}

void ref_vec4_add(const float* a, const float* b, float* out)
{
  Vec4f r = Vec4f(a[0], a[1], a[2], a[3]) + Vec4f(b[0], b[1], b[2], b[3]);
  for (int i = 0; i < 4; ++i) out[i] = r[i];
}

void ref_make_perspective(float fovy, float aspect, float znear, float zfar, float* out)
{
  store_mat(dp::math::makePerspective<float>(fovy, aspect, znear, zfar), out);
}

void ref_make_frustum(float l, float r, float b, float t, float n, float f, float* out)
{
  store_mat(dp::math::makeFrustum<float>(l, r, b, t, n, f), out);
}

void ref_make_ortho(float l, float r, float b, float t, float n, float f, float* out)
{
  store_mat(dp::math::makeOrtho<float>(l, r, b, t, n, f), out);
}

void ref_make_lookat(const float* eye, const float* center, const float* up, float* out)
{
  store_mat(dp::math::makeLookAt<float>(Vec3f(eye[0], eye[1], eye[2]), Vec3f(center[0], center[1], center[2]),
                                        Vec3f(up[0], up[1], up[2])),
            out);
}

// setMat(Mat44f, Quatf, trans, scale) (dp/math/Matmnt.h:527-537); q = (x, y, z, w).
void ref_set_mat_quat(const float* q, const float* trans, const float* scale, float* out)
{
  dp::math::Quatf quat(q[0], q[1], q[2], q[3]);
  Mat44f m;
  dp::math::setMat(m, quat, Vec3f(trans[0], trans[1], trans[2]), Vec3f(scale[0], scale[1], scale[2]));
  store_mat(m, out);
}

// ---- dp::util::BitArray ------------------------------------------------------
// BitArray(nbits); setBits(uint32 words, nbits); copy the 64-bit storage out.
void ref_bitarray_setbits(const uint32_t* src, size_t nbits, uint64_t* out_words)
{
  dp::util::BitArray b(nbits);
  b.setBits(src, nbits);
  size_t n = (nbits + 63) / 64;
  std::memcpy(out_words, b.getBits(), n * sizeof(uint64_t));
}

size_t ref_bitarray_count_leading_zeroes(const uint32_t* src, size_t nbits)
{
  dp::util::BitArray b(nbits);
  b.setBits(src, nbits);
  return b.countLeadingZeroes();
}

// traverseBits order: writes indices of set bits, returns the count.
size_t ref_bitarray_traverse(const uint32_t* src, size_t nbits, uint32_t* out_idx)
{
  dp::util::BitArray b(nbits);
  b.setBits(src, nbits);
  size_t count = 0;
  b.traverseBits([&](size_t i) { out_idx[count++] = static_cast<uint32_t>(i); });
  return count;
}

// ---- dp::transform::Tree -----------------------------------------------------
void* ref_tree_create() { return new RefTree(); }
void ref_tree_destroy(void* t) { delete static_cast<RefTree*>(t); }

// Returns 0 on success, -1 if the reference threw.
int ref_tree_add(void* t, uint32_t parent, const float* m, uint32_t* out_index)
{
  try {
    *out_index = static_cast<RefTree*>(t)->addTransform(parent, load_mat(m));
    return 0;
  } catch (std::exception const&) {
    return -1;
  }
}

int ref_tree_remove(void* t, uint32_t index)
{
  try {
    static_cast<RefTree*>(t)->removeTransform(index);
    return 0;
  } catch (std::exception const&) {
    return -1;
  }
}

void ref_tree_update_local(void* t, uint32_t index, const float* m)
{
  static_cast<RefTree*>(t)->updateLocalMatrix(index, load_mat(m));
}

// Runs Tree::compute and copies the event's dirty-world bits (one byte per index).
void ref_tree_compute(void* t, uint8_t* event_out)
{
  RefTree* tree = static_cast<RefTree*>(t);
  tree->compute(dp::math::cIdentity44f);
  std::memcpy(event_out, tree->m_lastEvent.data(), tree->m_lastEvent.size());
}

size_t ref_tree_count(void* t) { return static_cast<RefTree*>(t)->getTransformCount(); }

void ref_tree_world(void* t, float* out)
{
  RefTree* tree = static_cast<RefTree*>(t);
  const Mat44f* w = tree->getWorldMatrices();
  for (size_t i = 0; i < tree->getTransformCount(); ++i) store_mat(w[i], out + 16 * i);
}

size_t ref_tree_level_count(void* t) { return static_cast<RefTree*>(t)->levelCount(); }
size_t ref_tree_level_size(void* t, size_t l) { return static_cast<RefTree*>(t)->levelSize(l); }
void ref_tree_level_entries(void* t, size_t l, uint32_t* out) { static_cast<RefTree*>(t)->levelEntries(l, out); }
uint32_t ref_tree_node_level(void* t, uint32_t i) { return static_cast<RefTree*>(t)->nodeLevel(i); }
int ref_tree_is_free(void* t, uint32_t i) { return static_cast<RefTree*>(t)->isFree(i) ? 1 : 0; }

}  // extern "C"
