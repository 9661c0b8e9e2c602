// Copies of caller images between host memory and device rows on one stream.
//
// Page-locked host memory is copied directly.  Pageable memory goes through two page-locked
// chunks of kChunk bytes: the runtime's own pageable path measured 25 GB/s up and 17 GB/s
// down on MI355X, page-locked DMA 54-56 GB/s, and a host memcpy into a page-locked chunk
// about 29 GB/s.  Up, the CPU fills one chunk while the other one's DMA runs; down, the CPU
// drains one chunk while the DMA fills the other.
//
// One HostStaging belongs to one thread at a time.  A chunk whose DMA may still be pending
// (from this call or an earlier one, on any stream) is waited for before it is reused.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>

namespace ac::core::detail
{
    class HostStaging
    {
    public:
        static constexpr size_t kChunk = size_t(4) << 20;

        // Rows of `row` bytes from host `h` (pitch hpitch) to device `d` (pitch dpitch),
        // queued on `s`.  Pageable rows may be changed again once this returns; page-locked
        // rows are read by the DMA itself, so they must stay unchanged until `s` has passed it.
        bool toDevice(void* d, size_t dpitch, const void* h, size_t hpitch, size_t row, int rows, hipStream_t s)
        {
            const int per = perChunk(row, rows);
            if (per == 0 || pageLocked(h) || !ready())
                return check(hipMemcpy2DAsync(d, dpitch, h, hpitch, row, rows, hipMemcpyHostToDevice, s), "hipMemcpy2DAsync(H2D)");
            for (int r0 = 0, k = 0; r0 < rows; r0 += per, k ^= 1)
            {
                const int n = std::min(per, rows - r0);
                if (!wait(k)) return false;
                copyRows(buf_[k], row, static_cast<const char*>(h) + static_cast<size_t>(r0) * hpitch, hpitch, row, n);
                if (!check(hipMemcpy2DAsync(static_cast<char*>(d) + static_cast<size_t>(r0) * dpitch, dpitch, buf_[k], row, row, n,
                                            hipMemcpyHostToDevice, s), "hipMemcpy2DAsync(H2D)") ||
                    !check(hipEventRecord(ev_[k], s), "hipEventRecord"))
                    return false;
                pending_[k] = true;
            }
            return true;
        }

        // Rows from device `d` to host `h`, after everything queued on `s` before them.
        // Returns once the rows are in host memory.
        bool toHost(void* h, size_t hpitch, const void* d, size_t dpitch, size_t row, int rows, hipStream_t s)
        {
            const int per = perChunk(row, rows);
            if (per == 0 || pageLocked(h) || !ready())
                return check(hipMemcpy2DAsync(h, hpitch, d, dpitch, row, rows, hipMemcpyDeviceToHost, s), "hipMemcpy2DAsync(D2H)") &&
                       check(hipStreamSynchronize(s), "hipStreamSynchronize");
            const int chunks = (rows + per - 1) / per;
            auto issue = [&](const int c) {
                const int k = c & 1, r0 = c * per, n = std::min(per, rows - r0);
                if (!wait(k)) return false;
                if (!check(hipMemcpy2DAsync(buf_[k], row, static_cast<const char*>(d) + static_cast<size_t>(r0) * dpitch, dpitch, row, n,
                                            hipMemcpyDeviceToHost, s), "hipMemcpy2DAsync(D2H)") ||
                    !check(hipEventRecord(ev_[k], s), "hipEventRecord"))
                    return false;
                pending_[k] = true;
                return true;
            };
            for (int c = 0; c < std::min(chunks, 2); c++)
                if (!issue(c)) return false;
            for (int c = 0; c < chunks; c++)
            {
                const int r0 = c * per, n = std::min(per, rows - r0);
                if (!wait(c & 1)) return false;
                copyRows(static_cast<char*>(h) + static_cast<size_t>(r0) * hpitch, hpitch, buf_[c & 1], row, row, n);
                if (c + 2 < chunks && !issue(c + 2)) return false;
            }
            return true;
        }

        // Waits for pending DMAs and frees the chunks (with the device that made them current).
        void release() noexcept
        {
            for (int k = 0; k < 2; k++)
            {
                if (pending_[k]) (void)hipEventSynchronize(ev_[k]);
                pending_[k] = false;
                if (buf_[k]) (void)hipHostFree(buf_[k]);
                if (ev_[k]) (void)hipEventDestroy(ev_[k]);
                buf_[k] = nullptr;
                ev_[k] = nullptr;
            }
        }

        // the failed call and its error, after toDevice or toHost returned false
        const std::string& error() const noexcept { return error_; }

    private:
        static int perChunk(const size_t row, const int rows)
        {
            return row == 0 ? 0 : static_cast<int>(std::min<size_t>(kChunk / row, static_cast<size_t>(rows)));
        }

        static bool pageLocked(const void* p)
        {
            hipPointerAttribute_t a{};
            if (hipPointerGetAttributes(&a, p) != hipSuccess)
            {
                (void)hipGetLastError();   // pageable memory the runtime has never seen
                return false;
            }
            return a.type == hipMemoryTypeHost;
        }

        static void copyRows(void* d, const size_t dpitch, const void* s, const size_t spitch, const size_t row, const int rows)
        {
            if (dpitch == row && spitch == row)
            {
                std::memcpy(d, s, row * static_cast<size_t>(rows));
                return;
            }
            for (int r = 0; r < rows; r++)
                std::memcpy(static_cast<char*>(d) + static_cast<size_t>(r) * dpitch, static_cast<const char*>(s) + static_cast<size_t>(r) * spitch, row);
        }

        // allocates the chunks on first use; false (and the direct path) if it cannot
        bool ready()
        {
            for (int k = 0; k < 2; k++)
            {
                if (!buf_[k] && hipHostMalloc(&buf_[k], kChunk, 0) != hipSuccess)
                {
                    buf_[k] = nullptr;
                    (void)hipGetLastError();
                    return false;
                }
                if (!ev_[k] && hipEventCreateWithFlags(&ev_[k], hipEventDisableTiming) != hipSuccess)
                {
                    ev_[k] = nullptr;
                    (void)hipGetLastError();
                    return false;
                }
            }
            return true;
        }

        bool wait(const int k)
        {
            if (!pending_[k]) return true;
            pending_[k] = false;
            return check(hipEventSynchronize(ev_[k]), "hipEventSynchronize");
        }

        bool check(const hipError_t e, const char* what)
        {
            if (e == hipSuccess) return true;
            error_ = std::string(what) + ": " + hipGetErrorString(e);
            return false;
        }

        void* buf_[2] = {};
        hipEvent_t ev_[2] = {};
        bool pending_[2] = {};
        std::string error_;
    };
}
