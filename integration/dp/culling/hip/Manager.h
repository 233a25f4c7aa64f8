// dp/culling/hip/Manager.h -- the HIP (MI355X, gfx950) culling backend of
// nvpro-pipeline, beside dp/culling/cpu/Manager.h and dp/culling/opengl/Manager.h.
// Goes into the reference tree as dp/culling/hip/; selected by
// dp::culling::Mode::HIP in CullingImpl::CullingImpl
// (dp/sg/xbar/culling/src/CullingImpl.cpp:52-63).
#pragma once

#include <dp/culling/Config.h>
#include <dp/culling/Manager.h>

namespace dp
{
  namespace culling
  {
    namespace hip
    {
      class Manager
      {
      public:
        // A manager on the current HIP device; the caller owns it (CullingImpl
        // keeps it in a std::unique_ptr<dp::culling::Manager>, CullingImpl.h:113).
        DP_CULLING_API static dp::culling::Manager* create();
        // The same with the CPU backend's OBB caching reproduced exactly: box and
        // transform-index updates of objects already in a group take effect only
        // when the CPU backend would refresh its cache.
        DP_CULLING_API static dp::culling::Manager* createCpuCompatible();
      };
    } // namespace hip
  } // namespace culling
} // namespace dp
