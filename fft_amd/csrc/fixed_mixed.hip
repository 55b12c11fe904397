// fixed_mixed.hip -- compile-time plans, complex FP32: 1920 = 16 12 10, 3000 = 12 10 5 5.
// FixedC2C<type, N, threads per transform, transforms per block,
//          Sched<prefetch stage, min waves per SIMD, pair loads, prefetch first, LDS pad>,
//          variant (0 = default, the fastest measured), radices...>
#include "fixed.h"

namespace sfft {
namespace {
// 1920: 192 threads per transform give every stage one butterfly per thread
// (120, 160 and 192 butterflies); 160 threads leave the radix-10 stage with two
// butterflies on a fifth of the threads, measured 20 % slower.  One transform
// per block measured fastest on inputs the last-level cache does not hold
// (4.69 TB/s; two per block 4.45, four 4.17; 128 threads with two stages of
// two butterflies 4.57, 64 threads 4.43).
// default: non-temporal input loads and output stores (4.69 -> 4.81 TB/s;
// variant 8 is the same kernel without them)
const FixedC2C<float, 1920, 192, 1, Sched<0, 4, 0, 0, kLdsPadShift, 3>, 0, 16, 12, 10> c1920nt;
const FixedC2C<float, 1920, 192, 1, Sched<0, 4, 0>, 8, 16, 12, 10> c1920c;
const FixedC2C<float, 1920, 192, 2, Sched<0, 0, 0>, 1, 16, 12, 10> c1920b;
const FixedC2C<float, 1920, 192, 2, Sched<0, 4, 0>, 2, 16, 12, 10> c1920a;
const FixedC2C<float, 1920, 160, 2, Sched<0, 4, 0>, 3, 16, 12, 10> c1920d;
const FixedC2C<float, 1920, 128, 2, Sched<0, 0, 0>, 4, 16, 12, 10> c1920f;
const FixedC2C<float, 1920, 64, 4, Sched<0, 0, 0>, 5, 16, 12, 10> c1920h;
const FixedC2C<float, 1920, 192, 1, Sched<0, 4, 0, 0, 5>, 6, 16, 12, 10> c1920p5;
const FixedC2C<float, 1920, 192, 1, Sched<0, 4, 0, 0, kLdsNoPad>, 7, 16, 12, 10> c1920np;
// 3000: four stages (250, 300, 600, 600 butterflies on 300 threads: 10-12
// elements per thread in every stage) measured 2 % faster than three with a
// radix-25 stage that only 120 of 250 threads work in.  Its LDS exchanges have
// no power-of-two strides, and the usual one-in-16 padding caused more bank
// conflict cycles than LDS work (1.4x); unpadded it runs 4 % faster.
// default: non-temporal input loads and output stores (4.38 -> 4.52 TB/s;
// variant 7 is the same kernel without them)
// Other radix orders with the same accesses measured no faster: 10 12 5 5
// the same (4.54 vs 4.53 TB/s), 15 20 10 4.28, 20 15 10 4.26, 5 5 12 10 3.73.
const FixedC2C<float, 3000, 300, 1, Sched<0, 4, 0, 0, kLdsNoPad, 3>, 0, 12, 10, 5, 5> c3000nt;
const FixedC2C<float, 3000, 300, 1, Sched<0, 4, 0, 0, kLdsNoPad>, 7, 12, 10, 5, 5> c3000a;
const FixedC2C<float, 3000, 300, 1, Sched<0, 4, 0>, 1, 12, 10, 5, 5> c3000b;
const FixedC2C<float, 3000, 250, 1, Sched<0, 0, 0>, 2, 12, 10, 25> c3000c;
const FixedC2C<float, 3000, 250, 1, Sched<-1, 0, 0>, 3, 12, 10, 25> c3000d;
// fewer threads per transform, or more transforms per block, measured slower
// (150 x 2: 3.58 TB/s, 300 x 2: 3.77, 100 x 4: 3.27, against 4.42)
const FixedC2C<float, 3000, 300, 2, Sched<0, 4, 0, 0, kLdsNoPad>, 4, 12, 10, 5, 5> c3000f;
const FixedC2C<float, 3000, 300, 1, Sched<0, 4, 0, 0, 3>, 5, 12, 10, 5, 5> c3000p3;
const FixedC2C<float, 3000, 300, 1, Sched<0, 4, 0, 0, 5>, 6, 12, 10, 5, 5> c3000p5;
}  // namespace
}  // namespace sfft
