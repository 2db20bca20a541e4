// amx_internal.hpp -- shared definitions of the amat_amd library (not installed).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "amat_amd.h"

namespace amx {

// Failure carrying an amx_status code; caught at the C boundary.
struct Failure : std::runtime_error {
    int code;
    Failure(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const std::string& msg) { throw Failure(code, msg); }

void set_last_error(const std::string& msg);
void clear_last_error();

inline void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) {
        int code = (e == hipErrorOutOfMemory) ? AMX_OUT_OF_MEMORY : AMX_DEVICE_ERROR;
        fail(code, std::string(what) + ": " + hipGetErrorString(e));
    }
}
#define AMX_HIP(call) ::amx::hip_check((call), #call)

// Runs `body` and converts exceptions into a status code plus last-error text.
template <typename F> int guarded(const char* fn, F&& body)
{
    try {
        clear_last_error();
        body();
        return AMX_SUCCESS;
    } catch (const Failure& f) {
        set_last_error(std::string(fn) + ": " + f.what());
        return f.code;
    } catch (const std::bad_alloc&) {
        set_last_error(std::string(fn) + ": host allocation failed");
        return AMX_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        set_last_error(std::string(fn) + ": " + e.what());
        return AMX_INVALID_ARGUMENT;
    }
}

#define AMX_REQUIRE(cond, code, msg)                                                         \
    do {                                                                                     \
        if (!(cond)) ::amx::fail((code), (msg));                                             \
    } while (0)

// RAII device buffer (hipMalloc'd, coarse-grained).
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept
    {
        if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t n)
    {
        release();
        if (n == 0) return;
        AMX_HIP(hipMalloc(&p, n));
        bytes = n;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

// quadrature table of the FREE element kernels (amx_fem.cpp): nq points of
// [dN/dxi per node (3 each), weight], padded with zero-weight rows to rows_min
std::vector<double> free_quadrature(int kind, uint32_t ngt, uint32_t& nq, uint32_t rows_min);

}  // namespace amx

// Maps handle (host side), see amx_maps.cpp.
struct amx_maps {
    uint32_t rank = 0, size = 1;
    std::vector<uint32_t> rank_dof_counts;   // Allgather of owned counts
    std::vector<uint64_t> dof_scan;          // m_ulvLocalDofScan
    uint32_t n_elem = 0;
    std::vector<uint32_t> dpe;               // dofs per element
    uint32_t max_dpe = 0;                    // largest entry of dpe
    std::vector<uint64_t> eoff;              // prefix of dpe (n_elem + 1)
    std::vector<uint64_t> global_map;        // m_ulpMap flattened
    std::vector<uint32_t> local_map;         // m_uipLocalMap flattened
    uint32_t n_owned = 0, n_pre = 0, n_post = 0, n_total = 0;
    uint64_t owned_begin = 0, owned_end = 0, n_global = 0;
    std::vector<uint64_t> local2global;
    std::vector<uint32_t> localid2rank;
    std::vector<uint64_t> pre_gids, post_gids;
    std::vector<uint32_t> ghost_owner;       // owner of pre then post ghosts
    // exchange lists (include/maps.hpp:741-785)
    bool has_peer_lists = false;
    std::vector<uint32_t> send_counts, send_offsets, send_ids, send_rank_ids;
    std::vector<uint32_t> recv_counts, recv_offsets, recv_rank_ids;
    std::vector<uint32_t> independent, dependent;
    // boundary maps (include/maps.hpp:828-880)
    bool has_bdr = false;
    std::vector<uint32_t> bdr_map;
    std::vector<double> presval_map;
    std::vector<uint64_t> owned_cdofs;
    std::vector<double> owned_cvals;
    std::vector<uint64_t> owned_free;
    bool has_map = false;
};
