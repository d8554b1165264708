// Layout of the claim counters a persistent launch draws its blocks from (claim_block in
// kernels.hpp), shared by the kernels and by the host code that owns the counter memory.
//
// Device-scope atomics to one 64 B line are performed one after another: ~11.5 ns each on
// MI355X, so a single counter drawn on by 512 workgroups kept each claim in a queue for
// ~6 us, and 8 counters inside one line did no better.  A launch therefore spreads its
// claims over kClaimQueues counters, each on a 64 B line of its own (measured round trip of
// one claim with 512-1024 workgroups: ~0.6-0.7 us with 32 lines).
#pragma once

namespace fx {

constexpr int kClaimQueues = 32;
constexpr int kClaimLine = 16;                          // unsigned per 64 B line
constexpr int kClaimWords = kClaimQueues * kClaimLine;  // unsigned per launch (2 KiB)

}  // namespace fx
