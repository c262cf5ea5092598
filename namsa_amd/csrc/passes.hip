// Schedule registry: which compiled row / column schedule each transform
// length uses, and the C ABI to list and select them (the autotuner's knob).
#include <atomic>

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
// MI355X at the cache-sized probe group (3/4 of the 256 MB LLC of waves per
// launch, timed as back-to-back slice steps) among those whose round-off
// stays below 5e-5 of the CBED after 1000 vacuum slices.  The compact-twiddle
// (.ct) schedules are 3-7% faster but their power products add a coherent
// phase error per slice (1.6e-4 after 1000 slices), so they are opt-in.
//                                       128 192 256 384 512 768 1024 2048
static const int kDefaultRow[kNumSizes] = { 1,  0,  2,  0,  0,  0,  0,  0};
static const int kDefaultCol[kNumSizes] = { 0,  0,  2,  1, 10,  1,  1,  2};

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

int row_pass(int Nx, int op, const RowArgs& a, hipStream_t st) {
  const int i = size_index(Nx);
  if (i < 0) return fail(kErrUnsupported, "unsupported row length %d", Nx);
  init_selection();
  const SchedTable* t = table_for(Nx);
  const int sel = g_row_sel[i].load(std::memory_order_relaxed);
  return t->rows[sel].launch(op, a, st);
}

int col_pass(int Ny, int op, ColArgs& a, hipStream_t st) {
  const int i = size_index(Ny);
  if (i < 0) return fail(kErrUnsupported, "unsupported column length %d", Ny);
  init_selection();
  const SchedTable* t = table_for(Ny);
  const int sel = g_col_sel[i].load(std::memory_order_relaxed);
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

int namsa_sched_set(int n, int pass, int v) {
  const int i = size_index(n);
  if (i < 0 || v < 0 || v >= namsa_sched_count(n, pass))
    return fail(kErrArg, "namsa_sched_set: bad length %d, pass %d or schedule %d", n, pass, v);
  init_selection();
  (pass == PASS_ROW ? g_row_sel[i] : g_col_sel[i]).store(v);
  return kOk;
}

}  // extern "C"
