// self_front.hpp -- dpcull's own culling interface (include/dpcull/Manager.hpp)
// as the front end of dpcull::adapter::ManagerAdapter: the same method set as
// dp::culling::Manager (dp/culling/Manager.h:108-143), so the adapter that the
// reference tree compiles with dp::culling's types is compiled and run here.
#pragma once

#include "dpcull/Manager.hpp"
#include "dpcull_adapter.hpp"

struct SelfFront {
  typedef dpcull::culling::Manager Manager;
  typedef dpcull::culling::Object Object;
  typedef dpcull::culling::Group Group;
  typedef dpcull::culling::Result Result;
  typedef dpcull::culling::Payload Payload;
  typedef dpcull::Box3f Box3f;
  typedef dpcull::Mat44f Mat44f;
  static dpcull::Box3f toBack(Box3f const& b) { return b; }
  static Box3f toFront(dpcull::Box3f const& b) { return b; }
  static dpcull::Mat44f toBack(Mat44f const& m) { return m; }
};

typedef dpcull::adapter::ManagerAdapter<SelfFront> AdapterManager;
