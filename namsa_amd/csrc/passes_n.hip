// Compiled schedules of the row and column passes for one transform length.
// Built once per length with -DNAMSA_PASS_N=<n>; entry 0 is the base schedule.
// Names read: e<E> = elements per thread, r<radices>, c<C> = column tile,
// pf = next row/tile prefetched, nt = non-temporal wave traffic, ct = compact
// twiddles (one register pair per butterfly, higher powers formed in the stage),
// w<k> = register budget for at least k wavefronts per SIMD.
#include "passes_impl.hpp"

#ifndef NAMSA_PASS_N
#error "compile with -DNAMSA_PASS_N=<transform length>"
#endif

namespace namsa {

#if NAMSA_PASS_N == 128
using R0 = RowS<128, 8, Radices<8, 4, 4>, false, false>;
using R1 = RowS<128, 8, Radices<8, 4, 4>, true, false>;
using C0 = ColS<128, 8, Radices<8, 4, 4>, 16, false, false>;
using C1 = ColS<128, 8, Radices<8, 4, 4>, 8, true, false>;
using C2 = ColS<128, 8, Radices<8, 4, 4>, 16, false, false, false, false, 4, true, true>;
using R2 = RowS<128, 8, Radices<8, 4, 4>, false, false, false, 6>;
static const RowSchedEntry kRows[] = {{"e8.r844", launch_row<R0, R0>},
                                      {"e8.r844.pf", launch_row<R1, R0>},
                                      {"e8.r844.w6", launch_row<R2, R0>}};
static const ColSchedEntry kCols[] = {{"e8.r844.c16", launch_col<C0, C0>, 16},
                                      {"e8.r844.c8.pf", launch_col<C1, C0>, 8},
                                      {"e8.r844.c16.pg.sb.w4", launch_col<C2, C0>, 16}};
#elif NAMSA_PASS_N == 192
using R0 = RowS<192, 12, Radices<3, 4, 4, 4>, false, false>;
using R1 = RowS<192, 12, Radices<3, 4, 4, 4>, true, false>;
using C0 = ColS<192, 12, Radices<3, 4, 4, 4>, 16, false, false>;
using C1 = ColS<192, 12, Radices<3, 4, 4, 4>, 8, true, false>;
using R2 = RowS<192, 12, Radices<12, 4, 4>, false, false>;
using R3 = RowS<192, 12, Radices<12, 4, 4>, true, false>;
using C2 = ColS<192, 12, Radices<12, 4, 4>, 16, false, false>;
using C3 = ColS<192, 12, Radices<12, 4, 4>, 8, true, false>;
static const RowSchedEntry kRows[] = {{"e12.r3444", launch_row<R0, R0>},
                                      {"e12.r3444.pf", launch_row<R1, R0>},
                                      {"e12.r12.4.4", launch_row<R2, R0>},
                                      {"e12.r12.4.4.pf", launch_row<R3, R0>}};
static const ColSchedEntry kCols[] = {{"e12.r3444.c16", launch_col<C0, C0>, 16},
                                      {"e12.r3444.c8.pf", launch_col<C1, C0>, 8},
                                      {"e12.r12.4.4.c16", launch_col<C2, C0>, 16},
                                      {"e12.r12.4.4.c8.pf", launch_col<C3, C0>, 8}};
#elif NAMSA_PASS_N == 256
using R0 = RowS<256, 8, Radices<8, 8, 4>, false, false>;
using R1 = RowS<256, 8, Radices<8, 8, 4>, true, false>;
using R2 = RowS<256, 16, Radices<16, 16>, true, false>;
using R3 = RowS<256, 8, Radices<8, 8, 4>, false, false, false, 5>;
using C0 = ColS<256, 8, Radices<8, 8, 4>, 16, false, false>;
using C1 = ColS<256, 8, Radices<8, 8, 4>, 8, true, false>;
using C2 = ColS<256, 16, Radices<16, 16>, 16, true, false>;
static const RowSchedEntry kRows[] = {{"e8.r884", launch_row<R0, R0>},
                                      {"e8.r884.pf", launch_row<R1, R0>},
                                      {"e16.r16.16.pf", launch_row<R2, R0>},
                                      {"e8.r884.w5", launch_row<R3, R0>}};
static const ColSchedEntry kCols[] = {{"e8.r884.c16", launch_col<C0, C0>, 16},
                                      {"e8.r884.c8.pf", launch_col<C1, C0>, 8},
                                      {"e16.r16.16.c16.pf", launch_col<C2, C0>, 16}};
#elif NAMSA_PASS_N == 384
using R0 = RowS<384, 12, Radices<3, 2, 4, 4, 4>, false, false>;
using R1 = RowS<384, 12, Radices<3, 2, 4, 4, 4>, true, false>;
using C0 = ColS<384, 12, Radices<3, 2, 4, 4, 4>, 16, false, false>;
using C1 = ColS<384, 12, Radices<3, 2, 4, 4, 4>, 8, true, false>;
using R2 = RowS<384, 12, Radices<3, 2, 4, 4, 4>, false, false, true>;
using C2 = ColS<384, 12, Radices<3, 2, 4, 4, 4>, 8, true, false, false, true>;
using C3 = ColS<384, 12, Radices<3, 2, 4, 4, 4>, 8, true, false, true>;
using C4 = ColS<384, 12, Radices<3, 2, 4, 4, 4>, 8, false, false, false, false, 4, true, true>;
using C5 = ColS<384, 12, Radices<3, 2, 4, 4, 4>, 8, false, false, false, true, 4, true, true>;
using R3 = RowS<384, 12, Radices<12, 4, 4, 2>, false, false>;
using R4 = RowS<384, 12, Radices<12, 4, 4, 2>, false, false, true>;
using R5 = RowS<384, 12, Radices<6, 4, 4, 4>, false, false>;
using R6 = RowS<384, 12, Radices<6, 4, 4, 4>, false, false, true>;
using C6 = ColS<384, 12, Radices<12, 4, 4, 2>, 8, false, false, false, false, 4, true, true>;
using C7 = ColS<384, 12, Radices<12, 4, 4, 2>, 8, false, false, false, true, 4, true, true>;
using C8 = ColS<384, 12, Radices<6, 4, 4, 4>, 8, false, false, false, false, 4, true, true>;
using C9 = ColS<384, 12, Radices<6, 4, 4, 4>, 8, false, false, false, true, 4, true, true>;
static const RowSchedEntry kRows[] = {{"e12.r32444", launch_row<R0, R0>},
                                      {"e12.r32444.pf", launch_row<R1, R0>},
                                      {"e12.r32444.ct", launch_row<R2, R0>},
                                      {"e12.r12.4.4.2", launch_row<R3, R0>},
                                      {"e12.r12.4.4.2.ct", launch_row<R4, R0>},
                                      {"e12.r6444", launch_row<R5, R0>},
                                      {"e12.r6444.ct", launch_row<R6, R0>}};
static const ColSchedEntry kCols[] = {{"e12.r32444.c16", launch_col<C0, C0>, 16},
                                      {"e12.r32444.c8.pf", launch_col<C1, C0>, 8},
                                      {"e12.r32444.c8.pf.ct", launch_col<C2, C0>, 8},
                                      {"e12.r32444.c8.db.pf", launch_col<C3, C0>, 8},
                                      {"e12.r32444.c8.pg.sb.w4", launch_col<C4, C0>, 8},
                                      {"e12.r32444.c8.pg.sb.w4.ct", launch_col<C5, C0>, 8},
                                      {"e12.r12.4.4.2.c8.pg.sb.w4", launch_col<C6, C0>, 8},
                                      {"e12.r12.4.4.2.c8.pg.sb.w4.ct", launch_col<C7, C0>, 8},
                                      {"e12.r6444.c8.pg.sb.w4", launch_col<C8, C0>, 8},
                                      {"e12.r6444.c8.pg.sb.w4.ct", launch_col<C9, C0>, 8}};
#elif NAMSA_PASS_N == 512
using R0 = RowS<512, 8, Radices<8, 8, 8>, false, false>;
using R1 = RowS<512, 8, Radices<8, 8, 8>, true, false>;
using R2 = RowS<512, 8, Radices<8, 8, 8>, true, true>;
using R3 = RowS<512, 16, Radices<16, 16, 2>, false, false>;
using R4 = RowS<512, 16, Radices<16, 16, 2>, true, false>;
using C0 = ColS<512, 16, Radices<8, 8, 8>, 16, false, false>;
using C1 = ColS<512, 8, Radices<8, 8, 8>, 8, false, false>;
using C2 = ColS<512, 8, Radices<8, 8, 8>, 8, true, false>;
using C3 = ColS<512, 8, Radices<8, 8, 8>, 16, false, false>;
using C4 = ColS<512, 16, Radices<16, 16, 2>, 8, true, false>;
using C5 = ColS<512, 8, Radices<8, 8, 8>, 8, true, true>;
using C6 = ColS<512, 16, Radices<16, 16, 2>, 16, false, false>;
using C7 = ColS<512, 8, Radices<8, 8, 8>, 16, false, false, true>;
using C8 = ColS<512, 16, Radices<8, 8, 8>, 16, false, false, true>;
using C9 = ColS<512, 8, Radices<8, 8, 8>, 16, true, false, true>;
using C10 = ColS<512, 8, Radices<8, 8, 8>, 16, true, false>;
using C11 = ColS<512, 16, Radices<8, 8, 8>, 16, true, false, true>;
using C12 = ColS<512, 16, Radices<16, 16, 2>, 16, true, false>;
using C13 = ColS<512, 8, Radices<8, 8, 8>, 16, true, false, false, true>;
using C14 = ColS<512, 8, Radices<8, 8, 8>, 16, true, false, true, true>;
using C15 = ColS<512, 8, Radices<8, 8, 8>, 8, false, false, false, true, 6, true, true>;
using C16 = ColS<512, 8, Radices<8, 8, 8>, 8, false, false, false, false, 6, true, true>;
using C17 = ColS<512, 8, Radices<8, 8, 8>, 8, false, false, false, false, 6, true, true, 1, true>;
using R5 = RowS<512, 8, Radices<8, 8, 8>, false, false, true>;
using R6 = RowS<512, 8, Radices<8, 8, 8>, false, false, false, 5>;
static const RowSchedEntry kRows[] = {{"e8.r888", launch_row<R0, R0>},
                                      {"e8.r888.pf", launch_row<R1, R0>},
                                      {"e8.r888.pf.nt", launch_row<R2, R0>},
                                      {"e16.r16.16.2", launch_row<R3, R0>},
                                      {"e16.r16.16.2.pf", launch_row<R4, R0>},
                                      {"e8.r888.ct", launch_row<R5, R0>},
                                      {"e8.r888.w5", launch_row<R6, R0>}};
static const ColSchedEntry kCols[] = {{"e16.r888.c16", launch_col<C0, C0>, 16},
                                      {"e8.r888.c8", launch_col<C1, C0>, 8},
                                      {"e8.r888.c8.pf", launch_col<C2, C0>, 8},
                                      {"e8.r888.c16", launch_col<C3, C0>, 16},
                                      {"e16.r16.16.2.c8.pf", launch_col<C4, C0>, 8},
                                      {"e8.r888.c8.pf.nt", launch_col<C5, C0>, 8},
                                      {"e16.r16.16.2.c16", launch_col<C6, C0>, 16},
                                      {"e8.r888.c16.db", launch_col<C7, C0>, 16},
                                      {"e16.r888.c16.db", launch_col<C8, C0>, 16},
                                      {"e8.r888.c16.db.pf", launch_col<C9, C0>, 16},
                                      {"e8.r888.c16.pf", launch_col<C10, C0>, 16},
                                      {"e16.r888.c16.db.pf", launch_col<C11, C0>, 16},
                                      {"e16.r16.16.2.c16.pf", launch_col<C12, C0>, 16},
                                      {"e8.r888.c16.pf.ct", launch_col<C13, C0>, 16},
                                      {"e8.r888.c16.db.pf.ct", launch_col<C14, C0>, 16},
                                      {"e8.r888.c8.pg.sb.w6.ct", launch_col<C15, C0>, 8, "e8.r888.c8.lr.pg.sb.w6"},
                                      {"e8.r888.c8.pg.sb.w6", launch_col<C16, C0>, 8},
                                      {"e8.r888.c8.lr.pg.sb.w6", launch_col<C17, C0>, 8}};
#elif NAMSA_PASS_N == 768
using R0 = RowS<768, 12, Radices<3, 4, 4, 4, 4>, false, false>;
using R1 = RowS<768, 12, Radices<3, 4, 4, 4, 4>, true, false>;
using C0 = ColS<768, 12, Radices<3, 4, 4, 4, 4>, 8, false, false>;
using C1 = ColS<768, 12, Radices<3, 4, 4, 4, 4>, 8, true, false>;
using R2 = RowS<768, 12, Radices<3, 4, 4, 4, 4>, false, false, true>;
using C2 = ColS<768, 12, Radices<3, 4, 4, 4, 4>, 8, true, false, false, true>;
using C3 = ColS<768, 12, Radices<3, 4, 4, 4, 4>, 8, true, false, true>;
using C4 = ColS<768, 12, Radices<3, 4, 4, 4, 4>, 8, false, false, false, false, 4, true, true>;
using C5 = ColS<768, 12, Radices<3, 4, 4, 4, 4>, 8, false, false, false, true, 4, true, true>;
using C6 = ColS<768, 12, Radices<3, 4, 4, 4, 4>, 8, false, false, false, true, 4, false, true, 2>;
using R3 = RowS<768, 12, Radices<12, 4, 4, 4>, false, false>;
using R4 = RowS<768, 12, Radices<12, 4, 4, 4>, false, false, true>;
using C7 = ColS<768, 12, Radices<12, 4, 4, 4>, 8, false, false, false, true, 4, false, true, 2>;
using C8 = ColS<768, 12, Radices<12, 4, 4, 4>, 8, true, false, true>;
using C9 = ColS<768, 12, Radices<12, 4, 4, 4>, 8, false, false, false, true, 4, true, true>;
using C10 = ColS<768, 12, Radices<12, 4, 4, 4>, 8, false, false, false, false, 4, true, true>;
static const RowSchedEntry kRows[] = {{"e12.r34444", launch_row<R0, R0>},
                                      {"e12.r34444.pf", launch_row<R1, R0>},
                                      {"e12.r34444.ct", launch_row<R2, R0>},
                                      {"e12.r12.4.4.4", launch_row<R3, R0>},
                                      {"e12.r12.4.4.4.ct", launch_row<R4, R0>}};
static const ColSchedEntry kCols[] = {{"e12.r34444.c8", launch_col<C0, C0>, 8},
                                      {"e12.r34444.c8.pf", launch_col<C1, C0>, 8},
                                      {"e12.r34444.c8.pf.ct", launch_col<C2, C0>, 8},
                                      {"e12.r34444.c8.db.pf", launch_col<C3, C0>, 8},
                                      {"e12.r34444.c8.pg.sb.w4", launch_col<C4, C0>, 8},
                                      {"e12.r34444.c8.pg.sb.w4.ct", launch_col<C5, C0>, 8, "e12.r34444.c8.db.pf"},
                                      {"e12.r34444.c8.sb.w4.i2.ct", launch_col<C6, C0>, 8, "e12.r34444.c8.db.pf"},
                                      {"e12.r12.4.4.4.c8.sb.w4.i2.ct", launch_col<C7, C0>, 8, "e12.r12.4.4.4.c8.db.pf"},
                                      {"e12.r12.4.4.4.c8.db.pf", launch_col<C8, C0>, 8},
                                      {"e12.r12.4.4.4.c8.pg.sb.w4.ct", launch_col<C9, C0>, 8},
                                      {"e12.r12.4.4.4.c8.pg.sb.w4", launch_col<C10, C0>, 8}};
#elif NAMSA_PASS_N == 1024
using R0 = RowS<1024, 16, Radices<16, 8, 8>, false, false>;
using R1 = RowS<1024, 16, Radices<16, 8, 8>, true, false>;
using R2 = RowS<1024, 8, Radices<8, 8, 4, 4>, true, false>;
using R3 = RowS<1024, 16, Radices<16, 8, 8>, false, false, true>;
using R4 = RowS<1024, 16, Radices<16, 8, 8>, false, false, false, 3>;
using R5 = RowS<1024, 16, Radices<16, 8, 8>, false, false, false, 0, false, true>;
using C0 = ColS<1024, 16, Radices<16, 8, 8>, 8, false, false>;
using C1 = ColS<1024, 16, Radices<16, 8, 8>, 8, true, false>;
using C2 = ColS<1024, 8, Radices<8, 8, 4, 4>, 8, false, false>;
static const RowSchedEntry kRows[] = {{"e16.r16.8.8", launch_row<R0, R0>},
                                      {"e16.r16.8.8.pf", launch_row<R1, R0>},
                                      {"e8.r8844.pf", launch_row<R2, R0>},
                                      {"e16.r16.8.8.ct", launch_row<R3, R0>},
                                      {"e16.r16.8.8.w3", launch_row<R4, R0>},
                                      {"e16.r16.8.8.ls", launch_row<R5, R0>}};
using C3 = ColS<1024, 16, Radices<16, 8, 8>, 8, true, false, false, true>;
using C4 = ColS<1024, 16, Radices<16, 8, 8>, 8, false, false, false, false, 4, true, true>;
using C5 = ColS<1024, 16, Radices<16, 8, 8>, 8, false, false, false, true, 4, true, true>;
using C6 = ColS<1024, 16, Radices<16, 8, 8>, 8, false, false, false, true, 4, false, true, 4>;
using C7 = ColS<1024, 16, Radices<16, 8, 8>, 8, false, false, false, false, 4, true, true, 1, true>;
static const ColSchedEntry kCols[] = {{"e16.r16.8.8.c8", launch_col<C0, C0>, 8},
                                      {"e16.r16.8.8.c8.pf", launch_col<C1, C0>, 8},
                                      {"e8.r8844.c8", launch_col<C2, C0>, 8},
                                      {"e16.r16.8.8.c8.pf.ct", launch_col<C3, C0>, 8},
                                      {"e16.r16.8.8.c8.pg.sb.w4", launch_col<C4, C0>, 8},
                                      {"e16.r16.8.8.c8.pg.sb.w4.ct", launch_col<C5, C0>, 8},
                                      {"e16.r16.8.8.c8.sb.w4.i4.ct", launch_col<C6, C0>, 8, "e16.r16.8.8.c8.lr.pg.sb.w4"},
                                      {"e16.r16.8.8.c8.lr.pg.sb.w4", launch_col<C7, C0>, 8}};
#elif NAMSA_PASS_N == 2048
using R0 = RowS<2048, 16, Radices<16, 16, 8>, false, false>;
using R1 = RowS<2048, 16, Radices<16, 16, 8>, true, false>;
using C0 = ColS<2048, 16, Radices<16, 16, 8>, 4, false, false>;
using C1 = ColS<2048, 16, Radices<16, 16, 8>, 8, false, false>;
using C2 = ColS<2048, 16, Radices<16, 16, 8>, 4, true, false>;
using R2 = RowS<2048, 16, Radices<16, 16, 8>, false, false, true>;
using R3 = RowS<2048, 16, Radices<16, 16, 8>, false, false, false, 0, true>;
using R4 = RowS<2048, 16, Radices<16, 16, 8>, false, false, true, 0, true>;
using R5 = RowS<2048, 16, Radices<16, 16, 8>, false, false, false, 3, true, true>;
using C3 = ColS<2048, 16, Radices<16, 16, 8>, 4, true, false, false, true>;
using C4 = ColS<2048, 16, Radices<16, 16, 8>, 4, false, false, false, false, 4, true, true>;
using C5 = ColS<2048, 16, Radices<16, 16, 8>, 4, false, false, false, true, 4, true, true>;
using C6 = ColS<2048, 16, Radices<16, 16, 8>, 4, false, false, false, true, 4, false, true, 3>;
using C7 = ColS<2048, 16, Radices<16, 16, 8>, 4, false, false, false, false, 4, true, true, 1, true>;
static const RowSchedEntry kRows[] = {{"e16.r16.16.8", launch_row<R0, R0>},
                                      {"e16.r16.16.8.pf", launch_row<R1, R0>},
                                      {"e16.r16.16.8.ct", launch_row<R2, R0>},
                                      {"e16.r16.16.8.tr", launch_row<R3, R0>},
                                      {"e16.r16.16.8.tr.ct", launch_row<R4, R0>, "e16.r16.16.8.tr.ls"},
                                      {"e16.r16.16.8.tr.ls", launch_row<R5, R0>}};
static const ColSchedEntry kCols[] = {{"e16.r16.16.8.c4", launch_col<C0, C0>, 4},
                                      {"e16.r16.16.8.c8", launch_col<C1, C0>, 8},
                                      {"e16.r16.16.8.c4.pf", launch_col<C2, C0>, 4},
                                      {"e16.r16.16.8.c4.pf.ct", launch_col<C3, C0>, 4},
                                      {"e16.r16.16.8.c4.pg.sb.w4", launch_col<C4, C0>, 4},
                                      {"e16.r16.16.8.c4.pg.sb.w4.ct", launch_col<C5, C0>, 4},
                                      {"e16.r16.16.8.c4.sb.w4.i3.ct", launch_col<C6, C0>, 4, "e16.r16.16.8.c4.lr.pg.sb.w4"},
                                      {"e16.r16.16.8.c4.lr.pg.sb.w4", launch_col<C7, C0>, 4}};
#else
#error "no schedules for this transform length"
#endif

template <>
const SchedTable* sched_table<NAMSA_PASS_N>() {
  static const SchedTable table{kRows, static_cast<int>(sizeof(kRows) / sizeof(kRows[0])), kCols,
                                static_cast<int>(sizeof(kCols) / sizeof(kCols[0]))};
  return &table;
}

}  // namespace namsa
