// f2d_batch.hip -- member batching of kernel launches for ensembles.
//
// The reference advances one model per process (model.py:29-69).  An
// ensemble of many same-size members runs each operator of a stage as one
// launch for all of them: the launches of each member's stage are recorded
// (f2d_record_begin / f2d_record_end around the member's ordinary entry
// point calls, see launch_rec in f2d_common.hpp), the recordings are merged
// (f2d_launch_batch_create: kernel for kernel alike, argument structs into
// one device table per kernel) and replayed as one launch per kernel with
// blockIdx.y = member (f2d_launch_batch_run).  Each member's workgroups run
// the member's own arguments through the same code, so the results are its
// solo launches' bit for bit; the members must not share written arrays.
#include <memory>

#include "f2d_common.hpp"

namespace f2d {

Recording*& recording() {
  static thread_local Recording* r = nullptr;
  return r;
}

bool& recordable() {
  static thread_local bool ok = false;
  return ok;
}

}  // namespace f2d

using namespace f2d;

struct f2d_recording {
  Recording rec;
};

struct f2d_launch_batch {
  struct Op {
    const void* fn;
    dim3 grid, block;
    unsigned shmem;
    std::vector<unsigned char> a0;  // member 0's arguments (the plain argument slot)
    void* table;                    // device: the members' argument structs
  };
  int nm = 0;
  std::vector<Op> ops;
};

extern "C" {

int f2d_record_begin(void) {
  recordable() = true;  // this entry point itself runs while nothing records
  const int rc = guarded([&] {
    F2D_REQUIRE(!recording(), "a recording is already open on this thread");
    recording() = new Recording();
  });
  recordable() = false;
  return rc;
}

int f2d_record_end(f2d_recording** out) {
  recordable() = true;
  const int rc = guarded([&] {
    F2D_REQUIRE(recording(), "no recording is open on this thread");
    std::unique_ptr<Recording> r(recording());
    recording() = nullptr;
    F2D_REQUIRE(out, "out is NULL");
    auto* h = new f2d_recording();
    h->rec = std::move(*r);
    *out = h;
  });
  recordable() = false;
  return rc;
}

int f2d_recording_size(const f2d_recording* r) {
  return r ? (int)r->rec.launches.size() : -1;
}

int f2d_recording_destroy(f2d_recording* r) {
  delete r;
  return 0;
}

int f2d_launch_batch_create(f2d_recording* const* recs, int32_t n, f2d_launch_batch** out) {
  recordable() = true;
  const int rc = guarded([&] {
    F2D_REQUIRE(recs && out, "NULL argument");
    F2D_REQUIRE(n >= 1 && n <= 65535, "a batch of 1 .. 65535 members");
    for (int i = 0; i < n; i++) F2D_REQUIRE(recs[i], "a recording is NULL");
    const auto& r0 = recs[0]->rec.launches;
    for (int i = 1; i < n; i++) {
      const auto& ri = recs[i]->rec.launches;
      F2D_REQUIRE(ri.size() == r0.size(), "members recorded different numbers of launches");
      for (size_t p = 0; p < r0.size(); p++) {
        const LaunchRec &a = r0[p], &b = ri[p];
        F2D_REQUIRE(a.fn == b.fn && a.grid.x == b.grid.x && a.block.x == b.block.x &&
                        a.block.y == b.block.y && a.block.z == b.block.z &&
                        a.shmem == b.shmem && a.args.size() == b.args.size(),
                    "members recorded different launches (kernel, grid or block)");
      }
    }
    std::unique_ptr<f2d_launch_batch> bt(new f2d_launch_batch());
    bt->nm = n;
    try {
      for (size_t p = 0; p < r0.size(); p++) {
        const LaunchRec& a = r0[p];
        F2D_REQUIRE(a.grid.y == 1 && a.grid.z == 1, "a recorded launch has a 2-D grid");
        f2d_launch_batch::Op op{a.fn, a.grid, a.block, a.shmem, a.args, nullptr};
        const size_t sz = a.args.size();
        std::vector<unsigned char> host(sz * n);
        for (int i = 0; i < n; i++)
          std::memcpy(host.data() + sz * i, recs[i]->rec.launches[p].args.data(), sz);
        F2D_HIP(hipMalloc(&op.table, host.size()));
        bt->ops.push_back(std::move(op));
        F2D_HIP(hipMemcpy(bt->ops.back().table, host.data(), host.size(),
                          hipMemcpyHostToDevice));
      }
    } catch (...) {
      for (auto& op : bt->ops) (void)hipFree(op.table);
      throw;
    }
    *out = bt.release();
  });
  recordable() = false;
  return rc;
}

int f2d_launch_batch_run(f2d_launch_batch* bt, void* stream) {
  return guarded([&] {
    F2D_REQUIRE(bt, "batch is NULL");
    for (auto& op : bt->ops) {
      void* args[2] = {op.a0.data(), &op.table};
      F2D_HIP(hipLaunchKernel(op.fn, dim3(op.grid.x, (unsigned)bt->nm, 1), op.block, args,
                              op.shmem, S(stream)));
    }
  });
}

int f2d_launch_batch_destroy(f2d_launch_batch* bt) {
  return guarded([&] {
    if (!bt) return;
    F2D_HIP(hipDeviceSynchronize());
    for (auto& op : bt->ops) (void)hipFree(op.table);
    delete bt;
  });
}

}  // extern "C"
