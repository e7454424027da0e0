// Host-side copies of large caller arrays into pinned staging memory, split over a few
// threads.  One thread copies pageable memory at about 30 GB/s on the MI355X hosts,
// below what a pinned host-to-device copy sustains (about 57 GB/s); four threads copy at
// about 84 GB/s, so that staged copies keep the DMA engine busy.
#pragma once
#include <stddef.h>

namespace blsctx {

// dst[0..n) = src[0..n), split over the calling thread and the pool's workers; returns
// when every part has been copied.  Calls are serialized by the caller (the library
// lock); the workers are started on the first call that needs them and live until the
// process ends.
void host_copy(void* dst, const void* src, size_t n);

}  // namespace blsctx
