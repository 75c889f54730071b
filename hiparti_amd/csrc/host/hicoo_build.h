// Host pieces of the COO -> HiCOO conversion shared by the CPU conversion
// (hicoo.cpp) and the GPU one (device/convert_dev.hip).
#pragma once

#include <vector>

#include "internal.h"

namespace pti {

// Composite (superblock tuple, compressed Morton) key layout for the
// unique-order fast path: bits of mode m below lb[m] take part in the
// Morton interleave, bits from sk upward form the superblock tuple.
struct CompositeLayout {
    int N = 0, sk = 0;
    std::vector<int> lb, hb;
    int morton_bits = 0, total_bits = 0, max_lb = 0;
};
// False when the key needs more than 128 bits.
bool composite_layout(const std::vector<uint32_t> &mx, int sk, CompositeLayout &L);

// Per-mode superblock lists and nkiters (reference convert.c:378-421); h->kptr
// holds nonzero offsets of the superblock starts of the ordered tensor.
int set_kernel_scheduler(ptiSparseTensorHiCOO *h, const ptiSparseTensor *tsr, int sk);
// Balanced superblock partitions (reference convert.c:474-571).
int set_balanced_schedule(ptiSparseTensorHiCOO *h, int sk);
// cptr, and kptr rewritten to block offsets, from bptr (reference
// convert.c:728-835).
int set_chunks_from_blocks(ptiSparseTensorHiCOO *h);

}  // namespace pti
