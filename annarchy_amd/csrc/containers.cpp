// containers.cpp -- device-resident storage formats of a rate-coded projection,
// built on the host from the forward CSR view (rows = post ranks).
//
// Reference containers (host-side, include/*.hpp): CSRMatrix.hpp:142-193,
// ELLMatrix.hpp:290-349 (column-major for the GPU, padding INT_MAX),
// ELLRMatrix.hpp (row lengths), SELLMatrix.hpp:285-396 (slices without
// sorting), BSRMatrix.hpp:149-280, DenseMatrix.hpp:44.  The device layouts
// here are chosen for wave64 coalescing:
//   CSR       row_ptr int64 over all post ranks, col int32, w
//   ELL/ELLR  rows of the LIL only (empty rows are not stored), [k][row]
//   SELL      C = 64/TPR rows per slice, rows sorted by length (descending)
//             inside windows of sigma rows; slice width padded to a multiple
//             of TPR; entry (r, k) at off + (k/TPR)*64 + r*TPR + k%TPR
//   BSR       square tiles (2, 4 or 8), row-major inside a tile
//   DENSE     row-major n_rows x n_cols
//   CSR_PANEL CSR whose rows are split in column panels: panel-major order,
//             aux[q][r] = first entry of row r in panel q
//   SELL_PANEL the tiles of CSR_PANEL, each stored as sorted 64-row slices
//             with uint16 panel-relative columns
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <numeric>
#include <string>
#include <vector>

#include "internal.hpp"

namespace ahip {

namespace {

int upload(DevBuf& b, const void* src, size_t bytes, std::string& err) {
    b.release();
    if (bytes == 0) bytes = 8;  // keep a valid pointer for empty arrays
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        b.p = nullptr;
        err = std::string("hipMalloc failed: ") + hipGetErrorString(e);
        return AHIP_ENOMEM;
    }
    b.bytes = bytes;
    if (src) {
        e = hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            err = std::string("hipMemcpy failed: ") + hipGetErrorString(e);
            return AHIP_EHIP;
        }
    } else {
        // wait for the clear: later kernels run on a non-blocking stream
        e = hipMemset(b.p, 0, bytes);
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e != hipSuccess) {
            err = std::string("hipMemset failed: ") + hipGetErrorString(e);
            return AHIP_EHIP;
        }
    }
    return 0;
}

// Values in the network precision.
int upload_vals(DevBuf& b, const std::vector<double>& v, int32_t precision, std::string& err) {
    if (precision == AHIP_F64) return upload(b, v.data(), v.size() * sizeof(double), err);
    std::vector<float> f(v.size());
    for (size_t i = 0; i < v.size(); i++) f[i] = (float)v[i];
    return upload(b, f.data(), f.size() * sizeof(float), err);
}

inline double wval(const Proj& p, int64_t j) { return p.single_w ? p.w_const : p.h_w[j]; }

bool pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

}  // namespace

// Panel width for at most `pmax` columns per panel: the same number of
// panels, all of (about) equal width, so the last panel is not a sliver with
// too little work per tile.
// Compute units of the current device (256 on MI355X; that figure when no
// device answers, as on a build host).
static int32_t device_cus() {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) {
        (void)hipGetLastError();
        return 256;
    }
    return cus;
}

static int32_t equal_panels(int32_t n_cols, int32_t pmax) {
    if (n_cols <= 0 || pmax <= 0) return pmax;
    const int32_t np = (n_cols + pmax - 1) / pmax;
    return (n_cols + np - 1) / np;
}

int build_layout(const Proj& p, int32_t precision, const ahip_format_opts& o, RateLayout& L,
                 std::string& err) {
    const int32_t R = p.n_rows;
    const std::vector<int64_t>& rp = p.h_row_ptr;
    const std::vector<int32_t>& ci = p.h_col;
    L = RateLayout();
    L.format = o.format;
    int rc;
    switch (o.format) {
        case AHIP_FMT_CSR:
        case AHIP_FMT_CSR_VEC: {
            int lanes = o.lanes_per_row > 0 ? o.lanes_per_row : 64;
            if (!pow2(lanes) || lanes > 64) { err = "CSR lanes_per_row must be a power of two <= 64"; return AHIP_EINVAL; }
            if (o.format == AHIP_FMT_CSR_VEC && lanes < 4) { err = "csr_vec needs at least 4 lanes per row"; return AHIP_EINVAL; }
            L.lanes = lanes;
            std::vector<double> w(p.nnz);
            for (int64_t j = 0; j < p.nnz; j++) w[j] = wval(p, j);
            if ((rc = upload(L.ptr, rp.data(), rp.size() * sizeof(int64_t), err))) return rc;
            if ((rc = upload(L.idx, ci.data(), ci.size() * sizeof(int32_t), err))) return rc;
            if ((rc = upload_vals(L.val, w, precision, err))) return rc;
            L.stored = p.nnz;
            return 0;
        }
        case AHIP_FMT_ELL:
        case AHIP_FMT_ELLR: {
            std::vector<int32_t> rows;
            int32_t width = 0;
            for (int32_t r = 0; r < R; r++) {
                int64_t len = rp[r + 1] - rp[r];
                if (len > 0) { rows.push_back(r); width = std::max<int32_t>(width, (int32_t)len); }
            }
            const int32_t n = (int32_t)rows.size();
            std::vector<int32_t> col((size_t)n * width, INT32_MAX), rl(n);
            std::vector<double> val((size_t)n * width, 0.0);
            for (int32_t s = 0; s < n; s++) {
                const int32_t r = rows[s];
                rl[s] = (int32_t)(rp[r + 1] - rp[r]);
                for (int64_t j = rp[r], k = 0; j < rp[r + 1]; j++, k++) {
                    col[(size_t)k * n + s] = ci[j];
                    val[(size_t)k * n + s] = wval(p, j);
                }
            }
            L.width = width;
            L.n_units = n;
            L.stored = (int64_t)n * width;
            if ((rc = upload(L.perm, rows.data(), rows.size() * sizeof(int32_t), err))) return rc;
            if ((rc = upload(L.rl, rl.data(), rl.size() * sizeof(int32_t), err))) return rc;
            if ((rc = upload(L.idx, col.data(), col.size() * sizeof(int32_t), err))) return rc;
            if ((rc = upload_vals(L.val, val, precision, err))) return rc;
            return 0;
        }
        case AHIP_FMT_SELL: {
            int tpr = o.lanes_per_row > 0 ? o.lanes_per_row : 1;
            if (!pow2(tpr) || tpr > 64) { err = "SELL threads per row must be a power of two <= 64"; return AHIP_EINVAL; }
            const int C = 64 / tpr;
            int sigma = o.sigma > 0 ? o.sigma : 1;
            // row order: sort by length (descending, stable) inside sigma windows
            std::vector<int32_t> order(R);
            std::iota(order.begin(), order.end(), 0);
            if (sigma > 1) {
                for (int64_t w0 = 0; w0 < R; w0 += sigma) {
                    int64_t w1 = std::min<int64_t>(R, w0 + sigma);
                    std::stable_sort(order.begin() + w0, order.begin() + w1, [&](int32_t a, int32_t b) {
                        return (rp[a + 1] - rp[a]) > (rp[b + 1] - rp[b]);
                    });
                }
            }
            const int32_t ns = (R + C - 1) / C;
            std::vector<int64_t> off(ns + 1, 0);
            std::vector<int32_t> perm((size_t)ns * C, -1);
            for (int32_t s = 0; s < ns; s++) {
                int64_t wmax = 0;
                for (int r = 0; r < C; r++) {
                    int64_t q = (int64_t)s * C + r;
                    if (q < R) {
                        perm[q] = order[q];
                        wmax = std::max<int64_t>(wmax, rp[order[q] + 1] - rp[order[q]]);
                    }
                }
                int64_t wpad = (wmax + tpr - 1) / tpr * tpr;
                off[s + 1] = off[s] + wpad * C;
            }
            std::vector<int32_t> col(off[ns], 0);
            std::vector<double> val(off[ns], 0.0);
            for (int32_t s = 0; s < ns; s++) {
                for (int r = 0; r < C; r++) {
                    int64_t q = (int64_t)s * C + r;
                    if (q >= R) break;
                    const int32_t row = order[q];
                    for (int64_t j = rp[row], k = 0; j < rp[row + 1]; j++, k++) {
                        int64_t pos = off[s] + (k / tpr) * 64 + (int64_t)r * tpr + (k % tpr);
                        col[pos] = ci[j];
                        val[pos] = wval(p, j);
                    }
                }
            }
            L.lanes = tpr;
            L.chunk = C;
            L.sigma = sigma;
            L.n_units = ns;
            L.stored = off[ns];
            if ((rc = upload(L.ptr, off.data(), off.size() * sizeof(int64_t), err))) return rc;
            if ((rc = upload(L.perm, perm.data(), perm.size() * sizeof(int32_t), err))) return rc;
            if ((rc = upload(L.idx, col.data(), col.size() * sizeof(int32_t), err))) return rc;
            if ((rc = upload_vals(L.val, val, precision, err))) return rc;
            return 0;
        }
        case AHIP_FMT_BSR: {
            const int bs = o.chunk > 0 ? o.chunk : 4;
            if (bs != 2 && bs != 4 && bs != 8) { err = "BSR tile must be 2, 4 or 8"; return AHIP_EINVAL; }
            const int32_t nbr = (R + bs - 1) / bs;
            std::vector<int32_t> brp(nbr + 1, 0), bci;
            std::vector<double> val;
            for (int32_t b = 0; b < nbr; b++) {
                std::map<int32_t, std::vector<std::pair<int, double>>> tiles;
                for (int32_t r = b * bs; r < std::min(R, (b + 1) * bs); r++)
                    for (int64_t j = rp[r]; j < rp[r + 1]; j++)
                        tiles[ci[j] / bs].push_back({(r % bs) * bs + (ci[j] % bs), wval(p, j)});
                for (auto& kv : tiles) {
                    bci.push_back(kv.first);
                    size_t base = val.size();
                    val.resize(base + (size_t)bs * bs, 0.0);
                    for (auto& e : kv.second) val[base + e.first] = e.second;
                }
                brp[b + 1] = (int32_t)bci.size();
            }
            L.chunk = bs;
            L.n_units = nbr;
            L.stored = (int64_t)val.size();
            if ((rc = upload(L.ptr, brp.data(), brp.size() * sizeof(int32_t), err))) return rc;
            if ((rc = upload(L.idx, bci.data(), bci.size() * sizeof(int32_t), err))) return rc;
            if ((rc = upload_vals(L.val, val, precision, err))) return rc;
            return 0;
        }
        case AHIP_FMT_DENSE: {
            const size_t total = (size_t)R * p.n_cols;
            if (total > ((size_t)1 << 33)) { err = "dense layout too large"; return AHIP_EINVAL; }
            std::vector<double> val(total, 0.0);
            for (int32_t r = 0; r < R; r++)
                for (int64_t j = rp[r]; j < rp[r + 1]; j++) val[(size_t)r * p.n_cols + ci[j]] = wval(p, j);
            L.stored = (int64_t)total;
            if ((rc = upload_vals(L.val, val, precision, err))) return rc;
            return 0;
        }
        case AHIP_FMT_CSR_PANEL: {
            // tiles = (block of RB rows) x (panel of P columns); entries of a
            // tile are contiguous (rows ascending, columns ascending inside a
            // row); tile t = q * n_rb + rb, so consecutive workgroups share
            // the same slice of the pre-synaptic vector.
            int lanes = o.lanes_per_row > 0 ? o.lanes_per_row : 8;
            if (lanes != 4 && lanes != 8 && lanes != 16 && lanes != 32) { err = "tile lanes must be 4..32"; return AHIP_EINVAL; }
            const size_t es = precision == AHIP_F64 ? 8 : 4;
            int32_t pc = o.panel_cols > 0 ? o.panel_cols : (int32_t)(32768 / es);
            if ((size_t)pc * es > 64 * 1024) { err = "panel exceeds 64 KiB of LDS"; return AHIP_EINVAL; }
            pc = equal_panels(p.n_cols, pc);
            const int32_t rb_rows = o.chunk > 0 ? o.chunk : 512;
            if (rb_rows < 16 || rb_rows > 8192) { err = "tile rows must be 16..8192"; return AHIP_EINVAL; }
            const int32_t np = std::max<int32_t>(1, (p.n_cols + pc - 1) / pc);
            const int32_t nrb = (R + rb_rows - 1) / rb_rows;
            const int64_t ntiles = (int64_t)np * nrb;
            // count entries per (tile, local row)
            std::vector<int32_t> cnt((size_t)ntiles * rb_rows, 0);
            for (int32_t r = 0; r < R; r++) {
                const int32_t rb = r / rb_rows, lr = r % rb_rows;
                for (int64_t j = rp[r]; j < rp[r + 1]; j++) {
                    const int64_t t = (int64_t)(ci[j] / pc) * nrb + rb;
                    cnt[(size_t)t * rb_rows + lr]++;
                }
            }
            std::vector<int64_t> base(ntiles + 1, 0);
            std::vector<int32_t> roff((size_t)ntiles * (rb_rows + 1), 0);
            for (int64_t t = 0; t < ntiles; t++) {
                int32_t acc = 0;
                for (int32_t lr = 0; lr < rb_rows; lr++) {
                    roff[(size_t)t * (rb_rows + 1) + lr] = acc;
                    acc += cnt[(size_t)t * rb_rows + lr];
                }
                roff[(size_t)t * (rb_rows + 1) + rb_rows] = acc;
                base[t + 1] = base[t] + acc;
            }
            std::vector<int32_t> col(p.nnz);
            std::vector<double> val(p.nnz);
            std::fill(cnt.begin(), cnt.end(), 0);
            for (int32_t r = 0; r < R; r++) {
                const int32_t rb = r / rb_rows, lr = r % rb_rows;
                for (int64_t j = rp[r]; j < rp[r + 1]; j++) {
                    const int64_t t = (int64_t)(ci[j] / pc) * nrb + rb;
                    const int64_t pos = base[t] + roff[(size_t)t * (rb_rows + 1) + lr] + cnt[(size_t)t * rb_rows + lr]++;
                    col[pos] = ci[j];
                    val[pos] = wval(p, j);
                }
            }
            L.lanes = lanes;
            L.panel_cols = pc;
            L.chunk = rb_rows;
            L.n_units = np;
            L.width = nrb;
            L.stored = p.nnz;
            if ((rc = upload(L.ptr, base.data(), base.size() * sizeof(int64_t), err))) return rc;
            if ((rc = upload(L.aux, roff.data(), roff.size() * sizeof(int32_t), err))) return rc;
            if ((rc = upload(L.idx, col.data(), col.size() * sizeof(int32_t), err))) return rc;
            if ((rc = upload_vals(L.val, val, precision, err))) return rc;
            if ((rc = upload(L.perm, nullptr, (size_t)np * R * es, err))) return rc;  // partial sums
            return 0;
        }
        case AHIP_FMT_SELL_PANEL: {
            // tiles as for CSR_PANEL; inside a tile the RB rows are sorted by
            // their entry count in the tile (descending, stable) and packed
            // into slices of 64 lanes.  A slice has W steps of 4 entries per
            // lane; entry k of a row that owns lanes [b, b+g) of slice s sits
            // at off[s] + (e/g)*256 + 4*(b + e%g) + k%4 with e = k/4, so a
            // wave reads 4 synapses per lane with one 8-byte and one
            // 16/32-byte load.  Rows up to `cap` entries (twice the tile's
            // mean, at least 64) get one lane and W = longest row / 4; a
            // longer row gets g = ceil(len/cap) lanes (at most 64), W follows
            // from it and the rows after it in the slice get just enough
            // lanes to fit in W.  Columns are uint16, relative to the panel;
            // padding points at the zero slot P of the staged panel with
            // weight 0.  rowof[s][lane] = (g << 16) | row for the first lane
            // of a row, -1 for the others.
            const size_t es = precision == AHIP_F64 ? 8 : 4;
            // automatic: panels of up to 20479 columns (160 KiB of LDS in
            // FP64, one workgroup per CU), and tiles of at most 2048 rows,
            // as tall as lets every tile run in one round on the device's
            // workgroup slots (CUs x the panels one CU's LDS holds): every
            // slot busy, none waiting for a second round
            int32_t pc = o.panel_cols > 0 ? o.panel_cols : 20479;
            if (pc > 65535 || (size_t)(pc + 1) * es > 160 * 1024) { err = "panel exceeds 160 KiB of LDS or 65535 columns"; return AHIP_EINVAL; }
            pc = equal_panels(p.n_cols, pc);
            int32_t rb_rows = o.chunk;
            if (rb_rows <= 0) {
                const int64_t panels = std::max<int32_t>(1, (p.n_cols + pc - 1) / pc);
                const int64_t slots = (int64_t)device_cus() * std::max<int64_t>(1, (160 * 1024) / ((int64_t)(pc + 1) * es));
                const int64_t nrb = std::max<int64_t>(1, slots / panels);
                rb_rows = (int32_t)std::min<int64_t>(2048, std::max<int64_t>(256, ((R + nrb - 1) / nrb + 63) / 64 * 64));
            }
            if (rb_rows < 64 || rb_rows > 8192 || rb_rows % 64) { err = "tile rows must be a multiple of 64 in 64..8192"; return AHIP_EINVAL; }
            const int32_t steps = o.lanes_per_row > 0 ? o.lanes_per_row : 4;
            const int32_t waves = o.sigma > 1 ? o.sigma : 4;
            if (waves != 4 && waves != 8) { err = "sliced tiles: waves per workgroup must be 4 or 8"; return AHIP_EINVAL; }
            if (!(steps == 2 || steps == 4 || steps == 8 || (steps == 1 && waves == 4))) {
                err = "sliced tiles: steps in flight must be 2, 4 or 8 (1 with 4 waves)";
                return AHIP_EINVAL;
            }
            const int32_t np = std::max<int32_t>(1, (p.n_cols + pc - 1) / pc);
            // ceil(R / rb_rows) row blocks whose boundaries split the synapses
            // evenly (the psp ends with its heaviest tile; uneven in-degrees
            // would make equal-height blocks uneven), a block at most 65535
            // rows (the tile's row field) and 4 x rb_rows
            std::vector<int32_t> row0(1, 0);
            {
                const int32_t nrb0 = std::max<int32_t>(1, (R + rb_rows - 1) / rb_rows);
                const int64_t nnz = rp[R] - rp[0];
                const int32_t hcap = std::min<int32_t>(65535, 4 * rb_rows);
                int32_t r = 0;
                for (int32_t k = 1; k < nrb0 && r < R; k++) {
                    const int64_t target = rp[0] + (nnz * k + nrb0 - 1) / nrb0;
                    int32_t e = r;
                    while (e < R && rp[e] < target && e - r < hcap) e++;
                    // whole 64-row slices: the height rounded to the nearest
                    // multiple of 64 (equal in-degrees keep equal blocks)
                    const int32_t h = std::min<int32_t>({std::max<int32_t>(64, (e - r + 32) / 64 * 64), hcap, R - r});
                    row0.push_back(r + h);
                    r += h;
                }
                while (R - r > hcap) {
                    row0.push_back(r + hcap);
                    r += hcap;
                }
                if (r < R || row0.size() == 1) row0.push_back(R);
            }
            const int32_t nrb = (int32_t)row0.size() - 1;
            int32_t hmax = 1;
            std::vector<int32_t> rbof(R);
            for (int32_t b = 0; b < nrb; b++) {
                hmax = std::max(hmax, row0[b + 1] - row0[b]);
                for (int32_t r = row0[b]; r < row0[b + 1]; r++) rbof[r] = b;
            }
            const int64_t ntiles = (int64_t)np * nrb;
            std::vector<int32_t> cnt((size_t)ntiles * hmax, 0);
            for (int32_t r = 0; r < R; r++) {
                const int32_t rb = rbof[r], lr = r - row0[rb];
                for (int64_t j = rp[r]; j < rp[r + 1]; j++)
                    cnt[((size_t)(ci[j] / pc) * nrb + rb) * hmax + lr]++;
            }
            // slot[(t, lr)] = slice-in-tile << 12 | first lane << 6 | (lanes - 1)
            std::vector<int32_t> slot((size_t)ntiles * hmax, 0);
            std::vector<int32_t> tsl(ntiles + 1, 0);
            std::vector<int64_t> off(1, 0);
            std::vector<int32_t> rowof;
            std::vector<int32_t> order(hmax);
            for (int64_t t = 0; t < ntiles; t++) {
                const int32_t rb = (int32_t)(t % nrb);
                const int32_t nr = row0[rb + 1] - row0[rb];
                const int32_t* c = cnt.data() + (size_t)t * hmax;
                int64_t tile_nnz = 0;
                for (int32_t lr = 0; lr < nr; lr++) tile_nnz += c[lr];
                const int64_t cap = std::max<int64_t>(64, (2 * tile_nnz / nr + 3) / 4 * 4);
                std::iota(order.begin(), order.begin() + nr, 0);
                std::stable_sort(order.begin(), order.begin() + nr, [&](int32_t a, int32_t b) { return c[a] > c[b]; });
                int32_t ns = 0;
                for (int32_t i = 0; i < nr; ns++) {
                    const int64_t len0 = c[order[i]];
                    const int64_t g0 = len0 > cap ? std::min<int64_t>(64, (len0 + cap - 1) / cap) : 1;
                    const int64_t W = (len0 + 4 * g0 - 1) / (4 * g0);
                    int32_t used = 0;
                    rowof.resize(rowof.size() + 64, -1);
                    int32_t* ro = rowof.data() + rowof.size() - 64;
                    while (i < nr) {
                        const int64_t len = c[order[i]];
                        const int64_t g = used == 0 ? g0 : std::max<int64_t>(1, W > 0 ? (len + 4 * W - 1) / (4 * W) : 1);
                        if (used + g > 64) break;
                        ro[used] = (int32_t)((g << 16) | order[i]);
                        slot[(size_t)t * hmax + order[i]] = (ns << 12) | (used << 6) | (int32_t)(g - 1);
                        used += (int32_t)g;
                        i++;
                    }
                    off.push_back(off.back() + W * 256);
                }
                tsl[t + 1] = tsl[t] + ns;
            }
            const int64_t total = off.back();
            std::vector<uint16_t> col(total, (uint16_t)pc);
            std::vector<double> val(total, 0.0);
            std::fill(cnt.begin(), cnt.end(), 0);
            for (int32_t r = 0; r < R; r++) {
                const int32_t rb = rbof[r], lr = r - row0[rb];
                for (int64_t j = rp[r]; j < rp[r + 1]; j++) {
                    const int32_t q = ci[j] / pc;
                    const int64_t t = (int64_t)q * nrb + rb;
                    const int64_t k = cnt[(size_t)t * hmax + lr]++;
                    const int32_t sv = slot[(size_t)t * hmax + lr];
                    const int64_t gs = tsl[t] + (sv >> 12);
                    const int64_t b = (sv >> 6) & 63, g = (sv & 63) + 1;
                    const int64_t e = k / 4;
                    const int64_t pos = off[gs] + (e / g) * 256 + 4 * (b + e % g) + k % 4;
                    col[pos] = (uint16_t)(ci[j] - q * pc);
                    val[pos] = wval(p, j);
                }
            }
            L.lanes = steps;
            L.sigma = waves;
            L.panel_cols = pc;
            L.chunk = rb_rows;
            L.n_units = np;
            L.width = nrb;
            L.stored = total;
            if ((rc = upload(L.ptr, off.data(), off.size() * sizeof(int64_t), err))) return rc;
            if ((rc = upload(L.aux, rowof.data(), rowof.size() * sizeof(int32_t), err))) return rc;
            if ((rc = upload(L.tiles, tsl.data(), tsl.size() * sizeof(int32_t), err))) return rc;
            if ((rc = upload(L.rl, row0.data(), row0.size() * sizeof(int32_t), err))) return rc;  // row-block starts
            if ((rc = upload(L.idx, col.data(), col.size() * sizeof(uint16_t), err))) return rc;
            if ((rc = upload_vals(L.val, val, precision, err))) return rc;
            if ((rc = upload(L.perm, nullptr, (size_t)np * R * es, err))) return rc;  // partial sums
            return 0;
        }
        default:
            err = "unknown storage format";
            return AHIP_EINVAL;
    }
}

// Algorithmic bytes of one psp launch, independent of the layout (padding is
// not useful work): nnz*(s_w + 4) + (N_post+1)*4 + N_pre*s_r + 2*N_post*s_y.
double psp_bytes(const Proj& p, int32_t precision) {
    const double s = precision == AHIP_F64 ? 8.0 : 4.0;
    return (double)p.nnz * (s + 4.0) + ((double)p.n_rows + 1.0) * 4.0 + (double)p.n_cols * s +
           2.0 * (double)p.n_rows * s;
}

}  // namespace ahip
