// Memory pool over a few large chunks, for the serving loop's device buffers.
//
// hipMalloc takes 0.05-0.3 ms a call and a request needs up to eight buffers, so one
// hipMalloc per buffer would hold the scheduler thread (and with it the device) for
// milliseconds whenever many requests arrive at once.  Here chunks grow geometrically (the
// first `first` bytes, then as large as everything held so far, at most the cap); blocks are
// block-granular ranges of a chunk taken best-fit, and a freed range coalesces with its free
// neighbours inside the same chunk (two chunks that happen to be adjacent in the address
// space are never merged: each goes back to the backend whole).  A chunk returns to the
// backend when it is wholly free and an allocation fails, or when the pool is destroyed.
//
// Backend: bool alloc(void** p, size_t bytes) and void release(void* p).  Not thread-safe:
// one owner (the server's scheduler thread) calls it.
#pragma once

#include <algorithm>
#include <cstddef>
#include <map>
#include <unordered_map>

namespace achip {

template <typename Backend>
class ChunkPool {
public:
    explicit ChunkPool(Backend backend = Backend{}, size_t block = size_t(2) << 20, size_t first = size_t(256) << 20)
        : backend_(backend), block_(block), first_(round(first)), cap_(std::max(round(size_t(4) << 30), first_))
    {
    }
    ChunkPool(const ChunkPool&) = delete;
    ChunkPool& operator=(const ChunkPool&) = delete;
    ~ChunkPool() { trim(); }

    void set_chunk_cap(size_t cap) { cap_ = std::max(round(cap), first_); }
    size_t round(size_t v) const { return (std::max<size_t>(v, 1) + block_ - 1) / block_ * block_; }

    // A block of at least `bytes`; false when the backend cannot provide it.
    bool get(size_t bytes, void** out)
    {
        const size_t b = round(bytes);
        auto it = by_size_.lower_bound(b);
        if (it == by_size_.end()) {
            if (!grow(b)) return false;
            it = by_size_.lower_bound(b);
        }
        char* p = it->second;
        const size_t have = it->first;
        erase_free(p, have);
        if (have > b) insert_free(p + b, have - b);
        live_[p] = b;
        *out = p;
        return true;
    }

    // Back to the pool; pointers the pool did not hand out (and null) are ignored.
    void put(void* vp)
    {
        auto it = live_.find(vp);
        if (it == live_.end()) return;
        char* p = static_cast<char*>(vp);
        size_t n = it->second;
        live_.erase(it);
        auto next = free_.find(p + n);
        if (next != free_.end() && !chunks_.count(p + n)) {
            const size_t m = next->second;
            erase_free(p + n, m);
            n += m;
        }
        auto prev = free_.lower_bound(p);
        if (prev != free_.begin() && !chunks_.count(p)) {
            --prev;
            if (prev->first + prev->second == p) {
                char* q = prev->first;
                const size_t m = prev->second;
                erase_free(q, m);
                p = q;
                n += m;
            }
        }
        insert_free(p, n);
    }

    // Wholly free chunks back to the backend.
    void trim()
    {
        for (auto c = chunks_.begin(); c != chunks_.end();) {
            auto f = free_.find(c->first);
            if (f != free_.end() && f->second == c->second) {
                erase_free(c->first, c->second);
                backend_.release(c->first);
                held_ -= c->second;
                c = chunks_.erase(c);
            } else {
                ++c;
            }
        }
    }

    size_t held() const { return held_; }             // bytes in chunks
    size_t chunks() const { return chunks_.size(); }
    size_t free_ranges() const { return free_.size(); }
    size_t live_blocks() const { return live_.size(); }

private:
    bool grow(size_t b)
    {
        size_t n = std::max(b, std::min(cap_, std::max(first_, held_)));
        void* p = nullptr;
        if (!backend_.alloc(&p, n)) {
            trim();
            n = b;
            if (!backend_.alloc(&p, n)) return false;
        }
        chunks_[static_cast<char*>(p)] = n;
        held_ += n;
        insert_free(static_cast<char*>(p), n);
        return true;
    }
    void insert_free(char* p, size_t n)
    {
        free_[p] = n;
        by_size_.emplace(n, p);
    }
    void erase_free(char* p, size_t n)
    {
        free_.erase(p);
        for (auto r = by_size_.equal_range(n); r.first != r.second; ++r.first)
            if (r.first->second == p) {
                by_size_.erase(r.first);
                break;
            }
    }

    Backend backend_;
    size_t block_, first_, cap_;
    std::map<char*, size_t> chunks_;              // chunk base -> bytes
    std::map<char*, size_t> free_;                // free range start -> bytes (within one chunk)
    std::multimap<size_t, char*> by_size_;        // the same ranges by size, for best fit
    std::unordered_map<void*, size_t> live_;      // handed-out block -> bytes
    size_t held_ = 0;
};

} // namespace achip
