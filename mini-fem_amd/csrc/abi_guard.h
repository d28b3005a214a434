// Keeps C++ exceptions from crossing the C ABI: the body of an entry point that allocates
// host containers runs inside mf_guarded, which turns a failed host allocation into
// MF_ERR_NOMEM and any other exception into MF_ERR_STATE.
#pragma once

#include <new>

#include "minifem.h"

template <class F>
inline int mf_guarded(F &&body) noexcept
{
    try {
        return body();
    } catch (const std::bad_alloc &) {
        return MF_ERR_NOMEM;
    } catch (...) {
        return MF_ERR_STATE;
    }
}
