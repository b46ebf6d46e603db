// capi_guard.hpp — how the C ABI turns exceptions into error codes.
//
// Nothing throws across the boundary: an entry point runs its body through guarded(), which turns
// alien::Error (and any other exception) into a negative code plus a thread-local message, the way
// ALIEN's CHECK_FOR_CUDA_ERROR / checkAndThrowError turn CUDA errors into exceptions
// (source/EngineGpuKernels/Macros.cuh:15-59) one level up.
#pragma once

#include <new>
#include <string>

#include "../../include/alien_hip.h"
#include "worldset.hpp"

namespace alien {

// the message of the last failed call on this thread (alien_last_error)
std::string& last_error();

template <class F> int guarded(F&& f) {
    try {
        f();
        return ALIEN_OK;
    } catch (const Error& e) {
        last_error() = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        last_error() = "host allocation failed";
        return ALIEN_E_OOM;
    } catch (const std::exception& e) {
        last_error() = e.what();
        return ALIEN_E_STATE;
    }
}

}  // namespace alien
