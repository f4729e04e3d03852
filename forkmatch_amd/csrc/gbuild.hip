// The out-rows of a loaded .g graph built on the device, from the edge lines
// the device parser left in device memory: what build_graph (host.cpp) does
// on the host, with its checks and messages. The reference builds its CSR
// on the device too (sort, scan and csrToTrie, src/trie.cuh:800-951, with
// CUB's DeviceRadixSort::SortKeys and DeviceScan::ExclusiveScan); here the
// steps are a radix sort of (row, destination) keys, a scan over the
// de-duplicated keys and a scan over the row index, with AMD's rocPRIM
// device primitives (rocprim::radix_sort_keys, rocprim::exclusive_scan).
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "fm_internal.hpp"

namespace fm {

namespace {

constexpr uint32_t GB = 256;  // threads per block

#define GB_TRY(expr)                                                                          \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess) return fail(FM_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(_e)); \
    } while (0)

uint32_t grid_for(uint64_t n) { return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + GB - 1) / GB, 65536)); }

// First edge naming a vertex >= V (atomic minimum of its index) and the
// largest label value (one atomic per wave).
__global__ __launch_bounds__(GB) void gb_check_kernel(const uint32_t *src, const uint32_t *dst, const uint32_t *lab,
                                                      uint64_t n, uint32_t V, unsigned long long *first_bad,
                                                      uint32_t *max_label) {
    uint32_t m = 0;
    for (uint64_t i = blockIdx.x * (uint64_t)GB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * GB) {
        if (src[i] >= V || dst[i] >= V) atomicMin(first_bad, (unsigned long long)i);
        m = max(m, lab[i]);
    }
    for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o, 64));
    if ((threadIdx.x & 63u) == 0) atomicMax(max_label, m);
}

// present[v] = 1 for every label value v on an edge
__global__ __launch_bounds__(GB) void gb_mark_kernel(const uint32_t *lab, uint64_t n, uint32_t *present) {
    for (uint64_t i = blockIdx.x * (uint64_t)GB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * GB)
        present[lab[i]] = 1u;
}

// key = (row << vbits) | dst with row = dense label id * V + src
__global__ __launch_bounds__(GB) void gb_keys_kernel(const uint32_t *src, const uint32_t *dst, const uint32_t *lab,
                                                     uint64_t n, const uint32_t *lid, uint32_t V, uint32_t vbits,
                                                     unsigned long long *keys) {
    for (uint64_t i = blockIdx.x * (uint64_t)GB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * GB) {
        const unsigned long long row = (unsigned long long)lid[lab[i]] * V + src[i];
        keys[i] = (row << vbits) | dst[i];
    }
}

// first[i] = 1 where sorted key i differs from key i - 1 (duplicate lines
// collapse to one edge)
__global__ __launch_bounds__(GB) void gb_first_kernel(const unsigned long long *keys, uint64_t n, uint32_t *first) {
    for (uint64_t i = blockIdx.x * (uint64_t)GB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * GB)
        first[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

// the distinct keys' destinations into col, and the row lengths
__global__ __launch_bounds__(GB) void gb_scatter_kernel(const unsigned long long *keys, const uint32_t *first,
                                                        const uint32_t *pos, uint64_t n, uint32_t vbits,
                                                        uint32_t *col, uint32_t *len) {
    const unsigned long long vmask = (1ull << vbits) - 1ull;
    for (uint64_t i = blockIdx.x * (uint64_t)GB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * GB)
        if (first[i]) {
            col[pos[i]] = (uint32_t)(keys[i] & vmask);
            atomicAdd(&len[keys[i] >> vbits], 1u);
        }
}

// nonempty[r] = 1 for a row with edges (r < rows), 0 at r = rows
__global__ __launch_bounds__(GB) void gb_nonempty_kernel(const uint32_t *ptr, uint64_t rows, uint32_t *nonempty) {
    for (uint64_t r = blockIdx.x * (uint64_t)GB + threadIdx.x; r <= rows; r += (uint64_t)gridDim.x * GB)
        nonempty[r] = (r < rows && ptr[r + 1] > ptr[r]) ? 1u : 0u;
}

// roots[rpos[r]] = the vertex of every nonempty row r (labels in order, so
// each label's list is contiguous and ascending)
__global__ __launch_bounds__(GB) void gb_roots_kernel(const uint32_t *nonempty, const uint32_t *rpos, uint64_t rows,
                                                      uint32_t V, uint32_t *roots) {
    for (uint64_t r = blockIdx.x * (uint64_t)GB + threadIdx.x; r < rows; r += (uint64_t)gridDim.x * GB)
        if (nonempty[r]) roots[rpos[r]] = (uint32_t)(r % V);
}

// root_ptr[l] = rpos[l * V], l = 0..L
__global__ __launch_bounds__(GB) void gb_root_ptr_kernel(const uint32_t *rpos, uint32_t L, uint32_t V,
                                                         uint32_t *root_ptr) {
    const uint32_t l = blockIdx.x * GB + threadIdx.x;
    if (l <= L) root_ptr[l] = rpos[(uint64_t)l * V];
}

// row_sq[l] = sum over the rows of label l of the squared row length. Each
// thread sums a contiguous range of rows and adds per label it crosses; the
// terms are integers, so the sums are exact in any order.
__global__ __launch_bounds__(GB) void gb_row_sq_kernel(const uint32_t *ptr, uint64_t rows, uint32_t V,
                                                       uint64_t per_thread, double *row_sq) {
    const uint64_t t = blockIdx.x * (uint64_t)GB + threadIdx.x;
    const uint64_t r0 = t * per_thread;
    if (r0 >= rows) return;
    const uint64_t r1 = min(r0 + per_thread, rows);
    uint32_t l = (uint32_t)(r0 / V);
    uint64_t next = (uint64_t)(l + 1) * V;
    double acc = 0;
    uint32_t prev = ptr[r0];
    for (uint64_t r = r0; r < r1; r++) {
        if (r == next) {
            if (acc != 0) atomicAdd(&row_sq[l], acc);
            acc = 0;
            l++;
            next += V;
        }
        const uint32_t cur = ptr[r + 1];
        const double k = (double)(cur - prev);
        acc += k * k;
        prev = cur;
    }
    if (acc != 0) atomicAdd(&row_sq[l], acc);
}

// skip[b] = col[b << SKIP_SHIFT] for the blocks of the column array
__global__ __launch_bounds__(GB) void gb_skip_kernel(const uint32_t *col, uint64_t nskip, uint32_t *skip) {
    for (uint64_t b = blockIdx.x * (uint64_t)GB + threadIdx.x; b < nskip; b += (uint64_t)gridDim.x * GB)
        skip[b] = col[b << SKIP_SHIFT];
}

uint32_t bits_for(uint64_t maxv) {
    uint32_t b = 1;
    while (b < 64 && (maxv >> b) != 0) b++;
    return b;
}

// Device buffers freed on every return unless kept.
struct Bufs {
    std::vector<void *> v;
    ~Bufs() {
        for (void *p : v)
            if (p) (void)hipFree(p);
    }
    template <class T>
    int alloc(T **p, uint64_t count) {
        hipError_t e = hipMalloc((void **)p, std::max<uint64_t>(count, 1) * sizeof(T));
        if (e != hipSuccess) return fail(FM_ERR_DEVICE, "hipMalloc of %llu bytes: %s",
                                         (unsigned long long)(count * sizeof(T)), hipGetErrorString(e));
        v.push_back(*p);
        return FM_OK;
    }
    void keep(void *p) { std::replace(v.begin(), v.end(), p, (void *)nullptr); }
};

// out[i] = in[0] + ... + in[i - 1] (u32, exact: every sum here is a count
// of edges or rows below 2^32)
int exclusive_sum(Bufs &b, const uint32_t *in, uint32_t *out, uint64_t n) {
    size_t bytes = 0;
    GB_TRY(rocprim::exclusive_scan(nullptr, bytes, in, out, 0u, (size_t)n, rocprim::plus<uint32_t>()));
    char *tmp = nullptr;
    int rc = b.alloc(&tmp, std::max<size_t>(bytes, 1));
    if (rc) return rc;
    GB_TRY(rocprim::exclusive_scan(tmp, bytes, in, out, 0u, (size_t)n, rocprim::plus<uint32_t>()));
    return FM_OK;
}

}  // namespace

int gpu_build_graph(uint32_t V, uint64_t n, const uint32_t *d_edges, fm_graph *g, bool &built) {
    built = false;
    if (n == 0) return FM_OK;
    if (n >= (1ull << 32) - 1)
        return fail(FM_ERR_NOT_SUPPORTED, "graph has %llu directed edges; the CSR holds < 2^32",
                    (unsigned long long)n);
    const uint32_t *src = d_edges, *dst = d_edges + n, *lab = d_edges + 2 * n;
    Bufs b;
    int rc;
    // ---- vertex bounds, largest label ----
    unsigned long long *d_bad = nullptr;
    uint32_t *d_maxl = nullptr;
    if ((rc = b.alloc(&d_bad, 1)) || (rc = b.alloc(&d_maxl, 1))) return rc;
    GB_TRY(hipMemset(d_bad, 0xFF, 8));
    GB_TRY(hipMemset(d_maxl, 0, 4));
    hipLaunchKernelGGL(gb_check_kernel, dim3(grid_for(n)), dim3(GB), 0, 0, src, dst, lab, n, V, d_bad, d_maxl);
    GB_TRY(hipGetLastError());
    unsigned long long bad = 0;
    uint32_t maxl = 0;
    GB_TRY(hipMemcpy(&bad, d_bad, 8, hipMemcpyDeviceToHost));
    GB_TRY(hipMemcpy(&maxl, d_maxl, 4, hipMemcpyDeviceToHost));
    if (bad < n) {
        uint32_t s = 0, d = 0;
        GB_TRY(hipMemcpy(&s, src + bad, 4, hipMemcpyDeviceToHost));
        GB_TRY(hipMemcpy(&d, dst + bad, 4, hipMemcpyDeviceToHost));
        return fail(FM_ERR_DATA_BAD_BODY, "edge %llu (%u -> %u) names a vertex >= %u", bad, s, d, V);
    }
    if (maxl >= (1u << 24)) return FM_OK;
    // ---- dense label ids in ascending value order ----
    uint32_t *d_lid = nullptr;
    if ((rc = b.alloc(&d_lid, (uint64_t)maxl + 1))) return rc;
    GB_TRY(hipMemset(d_lid, 0, ((size_t)maxl + 1) * 4));
    hipLaunchKernelGGL(gb_mark_kernel, dim3(grid_for(n)), dim3(GB), 0, 0, lab, n, d_lid);
    GB_TRY(hipGetLastError());
    std::vector<uint32_t> lid((size_t)maxl + 1);
    GB_TRY(hipMemcpy(lid.data(), d_lid, lid.size() * 4, hipMemcpyDeviceToHost));
    std::vector<uint32_t> label_values;
    for (uint32_t v = 0; v <= maxl; v++)
        if (lid[v]) {
            lid[v] = (uint32_t)label_values.size();
            label_values.push_back(v);
        }
    GB_TRY(hipMemcpy(d_lid, lid.data(), lid.size() * 4, hipMemcpyHostToDevice));
    const uint32_t L = (uint32_t)label_values.size();
    const uint64_t rows = (uint64_t)L * V;
    if (rows > (1ull << 34))
        return fail(FM_ERR_NOT_SUPPORTED, "labels x vertices = %llu exceeds the dense row index limit 2^34",
                    (unsigned long long)rows);
    const uint32_t vbits = bits_for(V - 1), kbits = vbits + bits_for(rows - 1);
    if (kbits > 64) return FM_OK;
    // ---- (row, destination) keys, sorted ----
    unsigned long long *keys = nullptr, *sorted = nullptr;
    if ((rc = b.alloc(&keys, n)) || (rc = b.alloc(&sorted, n))) return rc;
    hipLaunchKernelGGL(gb_keys_kernel, dim3(grid_for(n)), dim3(GB), 0, 0, src, dst, lab, n, (const uint32_t *)d_lid,
                       V, vbits, keys);
    GB_TRY(hipGetLastError());
    {
        size_t bytes = 0;
        GB_TRY(rocprim::radix_sort_keys(nullptr, bytes, keys, sorted, (size_t)n, 0u, kbits));
        char *tmp = nullptr;
        if ((rc = b.alloc(&tmp, std::max<size_t>(bytes, 1)))) return rc;
        GB_TRY(rocprim::radix_sort_keys(tmp, bytes, keys, sorted, (size_t)n, 0u, kbits));
    }
    // ---- distinct edges: positions, destinations, row lengths ----
    uint32_t *first = nullptr, *pos = nullptr;
    if ((rc = b.alloc(&first, n)) || (rc = b.alloc(&pos, n))) return rc;
    hipLaunchKernelGGL(gb_first_kernel, dim3(grid_for(n)), dim3(GB), 0, 0, (const unsigned long long *)sorted, n,
                       first);
    GB_TRY(hipGetLastError());
    if ((rc = exclusive_sum(b, first, pos, n))) return rc;
    uint32_t last_pos = 0, last_first = 0;
    GB_TRY(hipMemcpy(&last_pos, pos + n - 1, 4, hipMemcpyDeviceToHost));
    GB_TRY(hipMemcpy(&last_first, first + n - 1, 4, hipMemcpyDeviceToHost));
    const uint64_t E = (uint64_t)last_pos + last_first;
    uint32_t *col = nullptr, *len = nullptr, *ptr = nullptr;
    if ((rc = b.alloc(&col, E + COL_PAD)) || (rc = b.alloc(&len, rows + 1)) || (rc = b.alloc(&ptr, rows + 1)))
        return rc;
    GB_TRY(hipMemset(col + E, 0, COL_PAD * 4));
    GB_TRY(hipMemset(len, 0, (rows + 1) * 4));
    hipLaunchKernelGGL(gb_scatter_kernel, dim3(grid_for(n)), dim3(GB), 0, 0, (const unsigned long long *)sorted,
                       (const uint32_t *)first, (const uint32_t *)pos, n, vbits, col, len);
    GB_TRY(hipGetLastError());
    if ((rc = exclusive_sum(b, len, ptr, rows + 1))) return rc;
    // ---- root lists: the nonempty rows, in row order ----
    uint32_t *nonempty = len, *rpos = nullptr;  // the lengths are no longer needed
    if ((rc = b.alloc(&rpos, rows + 1))) return rc;
    hipLaunchKernelGGL(gb_nonempty_kernel, dim3(grid_for(rows + 1)), dim3(GB), 0, 0, (const uint32_t *)ptr, rows,
                       nonempty);
    GB_TRY(hipGetLastError());
    if ((rc = exclusive_sum(b, nonempty, rpos, rows + 1))) return rc;
    uint32_t nroots = 0;
    GB_TRY(hipMemcpy(&nroots, rpos + rows, 4, hipMemcpyDeviceToHost));
    uint32_t *roots = nullptr, *root_ptr = nullptr;
    double *row_sq = nullptr;
    if ((rc = b.alloc(&roots, nroots)) || (rc = b.alloc(&root_ptr, (uint64_t)L + 1)) ||
        (rc = b.alloc(&row_sq, L)))
        return rc;
    hipLaunchKernelGGL(gb_roots_kernel, dim3(grid_for(rows)), dim3(GB), 0, 0, (const uint32_t *)nonempty,
                       (const uint32_t *)rpos, rows, V, roots);
    GB_TRY(hipGetLastError());
    hipLaunchKernelGGL(gb_root_ptr_kernel, dim3((L + GB) / GB), dim3(GB), 0, 0, (const uint32_t *)rpos, L, V,
                       root_ptr);
    GB_TRY(hipGetLastError());
    // ---- per-label squared row lengths (the planner's statistics) ----
    GB_TRY(hipMemset(row_sq, 0, (size_t)L * 8));
    {
        const uint64_t threads = (uint64_t)grid_for(rows) * GB, per = (rows + threads - 1) / threads;
        hipLaunchKernelGGL(gb_row_sq_kernel, dim3(grid_for(rows)), dim3(GB), 0, 0, (const uint32_t *)ptr, rows, V,
                           per, row_sq);
        GB_TRY(hipGetLastError());
    }
    // ---- skip index over the column array ----
    const uint64_t nskip = (E + (1u << SKIP_SHIFT) - 1) >> SKIP_SHIFT;
    uint32_t *skip = nullptr;
    if ((rc = b.alloc(&skip, nskip + COL_PAD))) return rc;
    GB_TRY(hipMemset(skip, 0, (nskip + COL_PAD) * 4));
    if (nskip) {
        hipLaunchKernelGGL(gb_skip_kernel, dim3(grid_for(nskip)), dim3(GB), 0, 0, (const uint32_t *)col, nskip, skip);
        GB_TRY(hipGetLastError());
    }
    // ---- the host copy ----
    HostGraph &h = g->h;
    h.V = V;
    h.label_values = label_values;
    h.out = Csr();
    h.in = Csr();
    h.out.ptr.resize(rows + 1);
    h.out.col.resize(E);
    h.out.roots.resize(nroots);
    h.out.root_ptr.resize((size_t)L + 1);
    h.out.row_sq.resize(L);
    GB_TRY(hipMemcpy(h.out.ptr.data(), ptr, (rows + 1) * 4, hipMemcpyDeviceToHost));
    GB_TRY(hipMemcpy(h.out.col.data(), col, E * 4, hipMemcpyDeviceToHost));
    GB_TRY(hipMemcpy(h.out.roots.data(), roots, (size_t)nroots * 4, hipMemcpyDeviceToHost));
    GB_TRY(hipMemcpy(h.out.root_ptr.data(), root_ptr, ((size_t)L + 1) * 4, hipMemcpyDeviceToHost));
    GB_TRY(hipMemcpy(h.out.row_sq.data(), row_sq, (size_t)L * 8, hipMemcpyDeviceToHost));
    h.symmetric = true;
    // ---- the device copy becomes the graph's upload ----
    int dev = 0;
    GB_TRY(hipGetDevice(&dev));
    b.keep(ptr);
    b.keep(col);
    b.keep(skip);
    b.keep(roots);
    g->d_ptr[0] = g->d_ptr[1] = ptr;
    g->d_col[0] = g->d_col[1] = col;
    g->d_skip = skip;
    g->d_roots[0] = g->d_roots[1] = roots;
    g->device = dev;
    g->device_bytes = ((rows + 1) + (E + COL_PAD) + (nskip + COL_PAD) + nroots) * 4;
    built = true;
    return FM_OK;
}

}  // namespace fm
