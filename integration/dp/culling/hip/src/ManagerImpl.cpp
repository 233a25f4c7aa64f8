// dp/culling/hip/src/ManagerImpl.cpp -- dp::culling::hip::Manager: the
// reference's culling interface (dp/culling/Manager.h:108-143) forwarded to the
// HIP backend (libdpcull_hip.so, include/dpcull/Manager.hpp) by
// dpcull::adapter::ManagerAdapter.  Compiled inside the reference tree
// (dp/culling/hip/CMakeLists.txt); the adapter itself is compiled and tested
// in this repository with dpcull's own interface as the front end.
#include <dp/culling/hip/Manager.h>

#include <dpcull_adapter.hpp>

namespace dp
{
  namespace culling
  {
    namespace hip
    {
      namespace
      {
        static_assert(sizeof(dp::math::Mat44f) == sizeof(dpcull::Mat44f) && sizeof(dp::math::Box3f) == sizeof(dpcull::Box3f),
                      "dp::math and dpcull value types have the same layout");

        // dp::culling's types as the adapter's front end.
        struct DpFront
        {
          typedef dp::culling::Manager Manager;
          typedef dp::culling::Object  Object;
          typedef dp::culling::Group   Group;
          typedef dp::culling::Result  Result;
          typedef dp::culling::Payload Payload;
          typedef dp::math::Box3f      Box3f;
          typedef dp::math::Mat44f     Mat44f;

          static dpcull::Box3f toBack(Box3f const& b)
          {
            dpcull::Box3f r;
            for (unsigned int c = 0; c < 3; ++c)
            {
              r.lower[c] = b.getLower()[c];
              r.upper[c] = b.getUpper()[c];
            }
            return r;
          }

          // The backend returns what ManagerBitSet::calculateBoundingBox returns,
          // Box3f(lower, upper) (ManagerBitSet.cpp:306).  Building Box3f(lower,
          // upper) from such a box again gives back the same bits: Boxnt(p0, p1)
          // (Boxnt.h:213-219) keeps p0 where p0 and p1 compare equal, which the
          // first construction already did.
          static Box3f toFront(dpcull::Box3f const& b)
          {
            return Box3f(dp::math::Vec3f(b.lower[0], b.lower[1], b.lower[2]),
                         dp::math::Vec3f(b.upper[0], b.upper[1], b.upper[2]));
          }

          static dpcull::Mat44f toBack(Mat44f const& m)
          {
            dpcull::Mat44f r;
            for (unsigned int i = 0; i < 4; ++i)
            {
              for (unsigned int j = 0; j < 4; ++j)
              {
                r[i][j] = m[i][j];
              }
            }
            return r;
          }
        };
      } // namespace

      dp::culling::Manager* Manager::create()
      {
        return new dpcull::adapter::ManagerAdapter<DpFront>();
      }

      dp::culling::Manager* Manager::createCpuCompatible()
      {
        dpcull::culling::hip::Options options;
        options.refreshOnObjectChange = false;
        return new dpcull::adapter::ManagerAdapter<DpFront>(options);
      }
    } // namespace hip
  } // namespace culling
} // namespace dp
