// A small persistent pool of host threads for per-chain host work (projected
// eigenproblems of many Lanczos runs at once).
#pragma once

#include <functional>

namespace dmrgx {

// f(i) for every i in [0, n), spread over at most 16 host threads (the
// caller's included); returns when all are done and rethrows the first
// exception any call threw.  Not re-entrant: f must not call parallel_for.
void parallel_for(int n, const std::function<void(int)>& f);

}  // namespace dmrgx
