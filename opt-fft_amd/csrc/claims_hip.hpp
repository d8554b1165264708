// ClaimSlotsT over the HIP runtime (claims.hpp).
#pragma once
#include <hip/hip_runtime.h>

#include "claims.hpp"

namespace fx {

struct HipClaimRT {
    using Stream = hipStream_t;
    using Event = hipEvent_t;
    using Error = hipError_t;
    static constexpr hipError_t ok = hipSuccess;
    // Returned when a plan has no capture counter left (kSlots captures already taken).
    static constexpr hipError_t busy = hipErrorStreamCaptureUnsupported;
    static hipError_t capturing(hipStream_t s, bool* yes)
    {
        hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
        const hipError_t e = hipStreamIsCapturing(s, &st);
        *yes = e == hipSuccess && st == hipStreamCaptureStatusActive;
        return e;
    }
    static hipError_t event_create(hipEvent_t* e)
    {
        return hipEventCreateWithFlags(e, hipEventDisableTiming);
    }
    static bool event_done(hipEvent_t e) { return hipEventQuery(e) == hipSuccess; }
    static hipError_t stream_wait(hipStream_t s, hipEvent_t e) { return hipStreamWaitEvent(s, e, 0); }
    static hipError_t event_record(hipEvent_t e, hipStream_t s) { return hipEventRecord(e, s); }
    static void event_destroy(hipEvent_t e) { (void)hipEventDestroy(e); }
};

using ClaimSlots = ClaimSlotsT<HipClaimRT>;

}  // namespace fx
