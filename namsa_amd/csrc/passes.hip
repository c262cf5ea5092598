// Schedule registry: which compiled row / column schedule each transform
// length uses, and the C ABI to list and select them (the autotuner's knob).
#include <atomic>
#include <cstring>

#include "namsa_hip.h"

#include "passes_impl.hpp"

namespace namsa {

#define NAMSA_FOR_EACH_N(X) X(0, 128) X(1, 192) X(2, 256) X(3, 384) X(4, 512) X(5, 768) X(6, 1024) X(7, 2048)
constexpr int kNumSizes = 8;

static int size_index(int n) {
  switch (n) {
#define X(i, n) case n: return i;
    NAMSA_FOR_EACH_N(X)
#undef X
    default:
      return -1;
  }
}

static const SchedTable* table_for(int n) {
  switch (n) {
#define X(i, n) case n: return sched_table<n>();
    NAMSA_FOR_EACH_N(X)
#undef X
    default:
      return nullptr;
  }
}

// Default schedule per length and pass: the fastest schedules measured on
// MI355X at the production probe group (timed as back-to-back slice steps).
// The row pass keeps its memory traffic under its arithmetic only with enough
// wavefronts per SIMD: at 512^2 the schedule holding every twiddle reaches
// that under a register budget (.w5), at 384^2 and 768^2 the radix-12 ones
// holding every twiddle do on their own, and at 1024^2 and 2048^2 the ones
// with split LDS padding (.ls: 191 -> 161 and 190 -> 168 VGPRs), 1-3%
// faster than the compact forms.  The fused driver switches compact
// schedules to full ones for thick runs (see precise_slices), where the
// compact forms' round-off would approach 1e-4.
// The column pass runs several workgroups per CU (.w<k>: register budget,
// one multi-workgroup launch of (strip, image) tasks; .sb: scalar-base
// addressing) at 128^2 and 384^2-2048^2, 4-24% faster per slice step than the
// persistent one-strip-per-workgroup schedules, which stay the defaults at
// 192^2 and 256^2.  Their compact forms keep the strip's propagator in
// registers for several images (.i<k>) at 1024^2-2048^2; the forms holding
// every twiddle read it at each use instead (.pg).  At 192^2, 384^2 and
// 768^2 the schedules whose first stage is a radix-12 prime-factor
// butterfly (three or four stages instead of four or five) run a slice step
// 11-23% faster than the radix-3/4 ones.  At 2048^2 the row step holds each
// transmission row across the group's images (.tr; 5-8% faster per slice
// step there, slower at smaller sizes, where the slice rows stay cached).
//                                       128 192 256 384 512 768 1024 2048
static const int kDefaultRow[kNumSizes] = { 1,  2,  0,  3,  6,  3,  5,  5};
static const int kDefaultCol[kNumSizes] = { 2,  2,  1,  7, 15,  9,  6,  6};

static std::atomic<int> g_row_sel[kNumSizes];
static std::atomic<int> g_col_sel[kNumSizes];

static void init_selection() {
  static const bool once = [] {
    for (int i = 0; i < kNumSizes; ++i) {
      g_row_sel[i].store(kDefaultRow[i]);
      g_col_sel[i].store(kDefaultCol[i]);
    }
    return true;
  }();
  (void)once;
}

bool supported(int n) { return size_index(n) >= 0; }

// The schedule holding every twiddle that a compact-twiddle ("<name>.ct")
// schedule is derived from (the base schedule 0 if it has no such partner),
// the partner its entry names explicitly (where the derived one is slower),
// or the schedule itself.
template <class Entry>
static int precise_partner(const Entry* list, int count, int sel) {
  if (const char* named = list[sel].precise) {
    for (int v = 0; v < count; ++v)
      if (std::strcmp(list[v].name, named) == 0) return v;
    return 0;
  }
  const char* name = list[sel].name;
  const size_t len = std::strlen(name);
  if (len < 3 || std::strcmp(name + len - 3, ".ct") != 0) return sel;
  for (int v = 0; v < count; ++v)
    if (std::strlen(list[v].name) == len - 3 && std::strncmp(list[v].name, name, len - 3) == 0) return v;
  return 0;
}

int row_pass(int Nx, int op, const RowArgs& a, hipStream_t st, bool precise) {
  const int i = size_index(Nx);
  if (i < 0) return fail(kErrUnsupported, "unsupported row length %d", Nx);
  init_selection();
  const SchedTable* t = table_for(Nx);
  int sel = g_row_sel[i].load(std::memory_order_relaxed);
  if (precise) sel = precise_partner(t->rows, t->n_rows, sel);
  return t->rows[sel].launch(op, a, st);
}

int col_pass(int Ny, int op, ColArgs& a, hipStream_t st, bool precise) {
  const int i = size_index(Ny);
  if (i < 0) return fail(kErrUnsupported, "unsupported column length %d", Ny);
  init_selection();
  const SchedTable* t = table_for(Ny);
  int sel = g_col_sel[i].load(std::memory_order_relaxed);
  if (precise) sel = precise_partner(t->cols, t->n_cols, sel);
  return t->cols[sel].launch(op, a, st);
}

}  // namespace namsa

using namespace namsa;

extern "C" {

int namsa_fft_supported(int n) { return supported(n) ? 1 : 0; }

int namsa_sched_count(int n, int pass) {
  const SchedTable* t = table_for(n);
  if (!t) return 0;
  return pass == PASS_ROW ? t->n_rows : (pass == PASS_COL ? t->n_cols : 0);
}

const char* namsa_sched_name(int n, int pass, int v) {
  const SchedTable* t = table_for(n);
  if (!t || v < 0 || v >= namsa_sched_count(n, pass)) return nullptr;
  return pass == PASS_ROW ? t->rows[v].name : t->cols[v].name;
}

int namsa_sched_get(int n, int pass) {
  const int i = size_index(n);
  if (i < 0 || (pass != PASS_ROW && pass != PASS_COL))
    return fail(kErrArg, "namsa_sched_get: bad length %d or pass %d", n, pass);
  init_selection();
  return pass == PASS_ROW ? g_row_sel[i].load() : g_col_sel[i].load();
}

int namsa_precise_slices(int n) { return supported(n) ? precise_slices(n) : -1; }

int namsa_sched_precise(int n, int pass, int v) {
  const SchedTable* t = table_for(n);
  if (!t || v < 0 || v >= namsa_sched_count(n, pass))
    return fail(kErrArg, "namsa_sched_precise: bad length %d, pass %d or schedule %d", n, pass, v);
  return pass == PASS_ROW ? precise_partner(t->rows, t->n_rows, v) : precise_partner(t->cols, t->n_cols, v);
}

int namsa_sched_set(int n, int pass, int v) {
  const int i = size_index(n);
  if (i < 0 || v < 0 || v >= namsa_sched_count(n, pass))
    return fail(kErrArg, "namsa_sched_set: bad length %d, pass %d or schedule %d", n, pass, v);
  init_selection();
  (pass == PASS_ROW ? g_row_sel[i] : g_col_sel[i]).store(v);
  return kOk;
}

}  // extern "C"
