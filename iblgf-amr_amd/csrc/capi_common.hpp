// Shared by the host and device halves of the C ABI: the tree handle and the
// status-code helpers (no C++ exception crosses the boundary).
#pragma once

#include "../../include/iblgf_gpu.h"

#include "host/ib.hpp"
#include "host/octree.hpp"

#include <stdexcept>
#include <string>

struct iblgf_tree
{
    iblgf_amd::Octree     tree;
    bool                  masks_built = false;
    bool                  subtract_non_leaf = false;  // the masks were built for it
    bool                  has_ib = false;
    iblgf_amd::IbGeometry ib;
};

namespace iblgf_amd {
namespace capi {

// Thrown by entry points for a case the device path does not build
// (IBLGF_ERR_UNSUPPORTED) and for a call in the wrong state (IBLGF_ERR_STATE).
struct Unsupported : std::runtime_error
{
    using std::runtime_error::runtime_error;
};
struct StateError : std::runtime_error
{
    using std::runtime_error::runtime_error;
};

// Records msg as the thread's last error (and in *ctx_error when given).
int fail(int code, const std::string& msg, std::string* ctx_error = nullptr);
// The calling thread's last error message.
const std::string& last_error();

// Runs f, mapping exceptions to status codes.
template<class F>
int guarded(std::string* ctx_error, F&& f)
{
    try
    {
        f();
        return IBLGF_OK;
    }
    catch (const Unsupported& e)
    {
        return fail(IBLGF_ERR_UNSUPPORTED, e.what(), ctx_error);
    }
    catch (const StateError& e)
    {
        return fail(IBLGF_ERR_STATE, e.what(), ctx_error);
    }
    catch (const std::invalid_argument& e)
    {
        return fail(IBLGF_ERR_ARG, e.what(), ctx_error);
    }
    catch (const std::length_error& e)
    {
        return fail(IBLGF_ERR_ARG, e.what(), ctx_error);
    }
    catch (const std::runtime_error& e)
    {
        const std::string m = e.what();
        if (m.rfind("HIP error", 0) == 0) return fail(IBLGF_ERR_DEVICE, m, ctx_error);
        if (m.find("unsupported") != std::string::npos) return fail(IBLGF_ERR_UNSUPPORTED, m, ctx_error);
        return fail(IBLGF_ERR_STATE, m, ctx_error);
    }
    catch (const std::exception& e)
    {
        return fail(IBLGF_ERR_INTERNAL, e.what(), ctx_error);
    }
    catch (...)
    {
        return fail(IBLGF_ERR_INTERNAL, "unknown error", ctx_error);
    }
}

}  // namespace capi
}  // namespace iblgf_amd
