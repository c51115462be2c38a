// Launch-shape autotuning: the first time a kernel family meets a shape,
// every applicable variant is timed on the device and the fastest is kept
// for the rest of the process (shared by all handles and batch instances).
#pragma once

#include <hip/hip_runtime.h>

#include <functional>
#include <string>

namespace chs {

// Cached choice for `key`, or -1 when the shape has not been tuned.
int tuned_choice(const std::string& key);
// The variant forced for `key` (chs_tune_force; -1 when none applies or it
// is not below nvar), else the cached choice (-1 when untuned).
int forced_or_tuned(const std::string& key, int nvar);

// Times launch(v) for v = 0 .. nvar-1 on stream s (one warm-up, best of three
// timed launches; launch returns false for a variant that does not apply) and
// caches the fastest.  Returns the choice, or -1 without timing when the
// stream is being captured into a graph or no variant applies.
int tune(const std::string& key, int nvar, hipStream_t s, const std::function<bool(int)>& launch);

// "key=choice (t0/t1/.. us)" lines of every tuned shape, for reports.
std::string tune_report();

// While variant >= 0, tune() returns `variant` (or -1 if it is >= nvar)
// without timing or caching for every key that starts with `prefix`, so
// tests can run each launch shape; variant < 0 restores tuning.
void tune_force(const std::string& prefix, int variant);

// Presets the choice for `key` as if it had been timed (e.g. from an earlier
// run's report), so a run under a profiler takes the same launch shapes.
void tune_set(const std::string& key, int choice);

}  // namespace chs
