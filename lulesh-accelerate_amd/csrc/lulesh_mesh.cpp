// lulesh_mesh.cpp -- host-side construction of the Sedov domain.
//
// The arithmetic that produces coordinates, volumes, masses and the initial
// time step keeps the reference's operation order, so the uploaded state is
// bit-identical to the reference's initial state.
#include "lulesh_mesh.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>

namespace lulesh {

GlibcRand::GlibcRand(unsigned seed)
{
    // glibc srandom_r for TYPE_3 (degree 31, separation 3): Park-Miller
    // seeding of 31 words, then 310 discarded outputs.  srand(0) seeds 1.
    int32_t r[31];
    if (seed == 0) {
        seed = 1;
    }
    r[0] = static_cast<int32_t>(seed);
    for (int i = 1; i < 31; ++i) {
        const int64_t hi = r[i - 1] / 127773;
        const int64_t lo = r[i - 1] % 127773;
        int64_t w = 16807 * lo - 2836 * hi;
        if (w < 0) {
            w += 2147483647;
        }
        r[i] = static_cast<int32_t>(w);
    }
    // ring_[j] holds the value 31 outputs back of slot j: r[3..30], r[0..2]
    for (int j = 0; j < 28; ++j) {
        ring_[j] = static_cast<uint32_t>(r[j + 3]);
    }
    for (int j = 28; j < 31; ++j) {
        ring_[j] = static_cast<uint32_t>(r[j - 28]);
    }
    pos_ = 0;
    for (int i = 0; i < 310; ++i) {
        next();
    }
}

int GlibcRand::next()
{
    // r[n] = r[n-31] + r[n-3] (mod 2^32); rand() returns r[n] >> 1
    const uint32_t v = ring_[pos_] + ring_[(pos_ + 28) % 31];
    ring_[pos_] = v;
    pos_ = (pos_ + 1) % 31;
    return static_cast<int>(v >> 1);
}

int region_rep(int ic_mode, int num_reg, int cost, int r)
{
    if (ic_mode == 1) {
        return 1;
    }
    if (r < num_reg / 2) {
        return 1;
    }
    if (r < num_reg - (num_reg + 15) / 20) {
        return 1 + cost;
    }
    return 10 * (1 + cost);
}

double hex_volume(const double x[8], const double y[8], const double z[8])
{
    auto tp = [](double x1, double y1, double z1, double x2, double y2,
                 double z2, double x3, double y3, double z3) {
        return x1 * (y2 * z3 - z2 * y3) + x2 * (z1 * y3 - y1 * z3) +
               x3 * (y1 * z2 - z1 * y2);
    };
    const double twelveth = 1.0 / 12.0;
    const double dx61 = x[6] - x[1], dy61 = y[6] - y[1], dz61 = z[6] - z[1];
    const double dx70 = x[7] - x[0], dy70 = y[7] - y[0], dz70 = z[7] - z[0];
    const double dx63 = x[6] - x[3], dy63 = y[6] - y[3], dz63 = z[6] - z[3];
    const double dx20 = x[2] - x[0], dy20 = y[2] - y[0], dz20 = z[2] - z[0];
    const double dx50 = x[5] - x[0], dy50 = y[5] - y[0], dz50 = z[5] - z[0];
    const double dx64 = x[6] - x[4], dy64 = y[6] - y[4], dz64 = z[6] - z[4];
    const double dx31 = x[3] - x[1], dy31 = y[3] - y[1], dz31 = z[3] - z[1];
    const double dx72 = x[7] - x[2], dy72 = y[7] - y[2], dz72 = z[7] - z[2];
    const double dx43 = x[4] - x[3], dy43 = y[4] - y[3], dz43 = z[4] - z[3];
    const double dx57 = x[5] - x[7], dy57 = y[5] - y[7], dz57 = z[5] - z[7];
    const double dx14 = x[1] - x[4], dy14 = y[1] - y[4], dz14 = z[1] - z[4];
    const double dx25 = x[2] - x[5], dy25 = y[2] - y[5], dz25 = z[2] - z[5];
    double v = tp(dx31 + dx72, dx63, dx20, dy31 + dy72, dy63, dy20,
                  dz31 + dz72, dz63, dz20) +
               tp(dx43 + dx57, dx64, dx70, dy43 + dy57, dy64, dy70,
                  dz43 + dz57, dz64, dz70) +
               tp(dx14 + dx25, dx61, dx50, dy14 + dy25, dy61, dy50,
                  dz14 + dz25, dz61, dz50);
    v *= twelveth;
    return v;
}

namespace {

// Region of every element: runs of 1..2048 elements drawn from the seeded
// generator, the region of a run chosen with weight (r+1)^balance and never
// equal to the previous run's region (lulesh-2.0.3/lulesh-init.cc:377-484).
void assign_regions(HostMesh &m, int balance)
{
    const int64_t ne = m.num_elem;
    const int nr = m.num_reg;
    m.reg_num_list.assign(static_cast<size_t>(ne), 1);
    if (nr == 1) {
        return;
    }
    std::vector<int> bin_end(static_cast<size_t>(nr));
    int cost_den = 0;
    for (int i = 0; i < nr; ++i) {
        cost_den += static_cast<int>(std::pow(double(i + 1), double(balance)));
        bin_end[static_cast<size_t>(i)] = cost_den;
    }
    // run-length classes: (threshold on rand()%1000, modulus, offset)
    struct RunClass { int below; int mod; int base; };
    static const RunClass classes[] = {
        {773, 15, 1},   {937, 16, 16},   {970, 32, 32},     {974, 64, 64},
        {978, 128, 128}, {981, 256, 256}, {1000, 1537, 512},
    };
    GlibcRand g(0);
    auto pick_region = [&]() {
        const int var = g.next() % cost_den;
        int i = 0;
        while (var >= bin_end[static_cast<size_t>(i)]) {
            ++i;
        }
        return (i % nr) + 1;
    };
    int last = -1;
    int64_t next = 0;
    while (next < ne) {
        int reg = pick_region();
        while (reg == last) {
            reg = pick_region();
        }
        const int bin = g.next() % 1000;
        int len = 0;
        for (const RunClass &c : classes) {
            if (bin < c.below) {
                len = g.next() % c.mod + c.base;
                break;
            }
        }
        const int64_t end = next + len;
        while (next < end && next < ne) {
            m.reg_num_list[static_cast<size_t>(next++)] = reg;
        }
        last = reg;
    }
}

}  // namespace

HostMesh build_mesh(const MeshConfig &cfg)
{
    if (cfg.nx < 1 || cfg.nx > 600) {
        throw std::invalid_argument("nx must be in [1, 600]");
    }
    if (cfg.ic_mode != 1 && cfg.ic_mode != 2) {
        throw std::invalid_argument("ic_mode must be 1 or 2");
    }
    if (cfg.ic_mode == 2 && (cfg.num_reg < 1 || cfg.num_reg > 4096 || cfg.cost < 0 || cfg.cost > 1000 || cfg.balance < 0 || cfg.balance > 16)) {
        throw std::invalid_argument("need 1 <= num_reg <= 4096, 0 <= cost <= 1000, 0 <= balance <= 16");
    }
    if (!(cfg.efactor > 0.0) || !(cfg.efactor < 1.0e6)) {
        throw std::invalid_argument("efactor must be in (0, 1e6)");
    }
    HostMesh m;
    const int nx = cfg.nx, en = nx + 1;
    m.nx = nx;
    m.edge_nodes = en;
    m.num_elem = int64_t(nx) * nx * nx;
    m.num_node = int64_t(en) * en * en;
    const size_t ne = static_cast<size_t>(m.num_elem);
    const size_t nn = static_cast<size_t>(m.num_node);

    // nodal coordinates (reference/C/LULESH-OMP.cc:2996-3017)
    m.x.resize(nn);
    m.y.resize(nn);
    m.z.resize(nn);
    std::vector<double> coord(static_cast<size_t>(en));
    coord[0] = 0.0;
    for (int c = 1; c < en; ++c) {
        coord[static_cast<size_t>(c)] = 1.125 * double(c) / double(nx);
    }
    for (int k = 0; k < en; ++k) {
        for (int j = 0; j < en; ++j) {
            for (int i = 0; i < en; ++i) {
                const size_t n = (size_t(k) * en + j) * en + i;
                m.x[n] = coord[static_cast<size_t>(i)];
                m.y[n] = coord[static_cast<size_t>(j)];
                m.z[n] = coord[static_cast<size_t>(k)];
            }
        }
    }

    // element volumes, masses and nodal masses, in element order
    // (reference/C/LULESH-OMP.cc:3089-3109)
    m.volo.resize(ne);
    m.nodal_mass.assign(nn, 0.0);
    for (int k = 0; k < nx; ++k) {
        for (int j = 0; j < nx; ++j) {
            for (int i = 0; i < nx; ++i) {
                const size_t el = (size_t(k) * nx + j) * nx + i;
                const size_t n0 = (size_t(k) * en + j) * en + i;
                const size_t nodes[8] = {
                    n0, n0 + 1, n0 + en + 1, n0 + en,
                    n0 + size_t(en) * en, n0 + size_t(en) * en + 1,
                    n0 + size_t(en) * en + en + 1, n0 + size_t(en) * en + en};
                double xl[8], yl[8], zl[8];
                for (int c = 0; c < 8; ++c) {
                    xl[c] = m.x[nodes[c]];
                    yl[c] = m.y[nodes[c]];
                    zl[c] = m.z[nodes[c]];
                }
                const double vol = hex_volume(xl, yl, zl);
                m.volo[el] = vol;
                for (int c = 0; c < 8; ++c) {
                    m.nodal_mass[nodes[c]] += vol / 8.0;
                }
            }
        }
    }

    // initial energy and time step
    m.e.assign(ne, 0.0);
    if (cfg.ic_mode == 1) {
        // reference/C/LULESH-OMP.cc:3052-3053, :3112
        m.e[0] = 3.948746e+7 * cfg.efactor;
        m.deltatime0 = 1.0e-7;
        m.dtfixed = -1.0e-7;
    } else {
        // reference/C/lulesh-2.0.3/lulesh-init.cc:140, :167-182; an ensemble
        // member's energy factor multiplies the scaled energy last (x 1.0 is
        // exact, so the single run is unchanged)
        const double ebase = 3.948746e+7;
        const double scale = double(nx * 1) / 45.0;
        const double einit = ebase * scale * scale * scale * cfg.efactor;
        m.e[0] = einit;
        m.deltatime0 = (.5 * std::cbrt(m.volo[0])) / std::sqrt(2.0 * einit);
        m.dtfixed = -1.0e-6;
    }

    // regions and the EOS evaluation order
    m.num_reg = (cfg.ic_mode == 1) ? 1 : cfg.num_reg;
    assign_regions(m, cfg.balance);
    m.reg_elem_size.assign(static_cast<size_t>(m.num_reg), 0);
    for (size_t i = 0; i < ne; ++i) {
        m.reg_elem_size[static_cast<size_t>(m.reg_num_list[i] - 1)]++;
    }
    m.region_offset.assign(static_cast<size_t>(m.num_reg) + 1, 0);
    for (int r = 0; r < m.num_reg; ++r) {
        m.region_offset[static_cast<size_t>(r) + 1] =
            m.region_offset[static_cast<size_t>(r)] + m.reg_elem_size[static_cast<size_t>(r)];
    }
    m.eos_order.resize(ne);
    m.eos_rep.resize(ne);
    m.elem_rep.resize(ne);
    std::vector<int32_t> fill(m.region_offset.begin(), m.region_offset.end() - 1);
    for (size_t i = 0; i < ne; ++i) {
        const int r = m.reg_num_list[i] - 1;
        const int32_t slot = fill[static_cast<size_t>(r)]++;
        m.eos_order[static_cast<size_t>(slot)] = static_cast<int32_t>(i);
        m.eos_rep[static_cast<size_t>(slot)] = static_cast<uint16_t>(
            region_rep(cfg.ic_mode, m.num_reg, cfg.cost, r));
        m.elem_rep[i] = m.eos_rep[static_cast<size_t>(slot)];
    }
    m.eos_blocks = eos_block_order(m.elem_rep);
    return m;
}

// A block's cost is its EOS passes: one for the elements' own evaluations,
// plus the evaluations beyond each element's first shared over the block's
// threads, ceil(sum(rep - 1) / kEosBlockElems).  With LULESH 2.0's regions
// the costs are far from even: at 150^3 (cost 1, 11 regions) 8% of the
// blocks lie in runs of the 20-repetition region and take 20 passes, 40%
// take 2, the mean is 4.6.  A costly block started late runs on after the
// rest of the launch has finished; costly blocks all started together leave
// the memory idle, and the cheap ones left for the end leave the ALUs idle.
// The hardware deals slot t of a launch to XCD t % 8.  So the costly blocks
// (above kHeavyCost times the mean), costliest first, are dealt round the
// XCDs, and each XCD spreads its share evenly over the first kHeavySpan of
// its slots; the rest fill its other slots in mesh order, each XCD taking
// one contiguous range of them, so the neighbour planes one element layer
// shares with the next stay in that XCD's L2.  Measured on the Q/EOS kernel
// against mesh order alone: 90^3 52.0 -> 38.2 us, 120^3 90.3 -> 75.3 us,
// 150^3 146.3 -> 140.5 us (float64 with fused multiply-adds); mixed
// precision 120^3 64.8 -> 54.8 us.
namespace {

constexpr double kHeavyCost = 1.0, kHeavySpan = 0.75;

}  // namespace

std::vector<int64_t> eos_block_passes(const std::vector<uint16_t> &rep)
{
    const size_t ne = rep.size(), be = size_t(kEosBlockElems);
    const size_t g = (ne + be - 1) / be;
    std::vector<int64_t> passes(g);
    for (size_t b = 0; b < g; ++b) {
        int64_t extra = 0;
        for (size_t i = b * be; i < std::min(ne, (b + 1) * be); ++i) {
            extra += int64_t(rep[i]) - 1;
        }
        passes[b] = 1 + (extra + int64_t(be) - 1) / int64_t(be);
    }
    return passes;
}

std::vector<int32_t> eos_block_order(const std::vector<uint16_t> &rep)
{
    return spread_costly_blocks(eos_block_passes(rep));
}

std::vector<int32_t> spread_costly_blocks(const std::vector<int64_t> &passes)
{
    const size_t g = passes.size();
    int64_t sum = 0;
    for (int64_t p : passes) sum += p;
    const double heavy_min = kHeavyCost * double(sum) / double(std::max<size_t>(g, 1));
    std::vector<int32_t> heavy, rest;
    for (size_t b = 0; b < g; ++b) {
        (double(passes[b]) > heavy_min ? heavy : rest).push_back(int32_t(b));
    }
    std::stable_sort(heavy.begin(), heavy.end(),
                     [&](int32_t a, int32_t b) { return passes[size_t(a)] > passes[size_t(b)]; });
    // slot t = 8 r + x is XCD x's r-th; XCD x takes heavy[x], heavy[x + 8],
    // ... at evenly spaced r over the first kHeavySpan of its n_x slots, and
    // one contiguous range of the rest in its other slots
    const size_t h = heavy.size();
    std::vector<char> heavy_slot(g, 0);
    std::vector<int32_t> order(g);
    size_t start[8] = {};
    for (size_t x = 0; x < 8 && x < g; ++x) {
        const size_t n_x = (g - x + 7) / 8;
        const size_t h_x = (h > x) ? (h - x + 7) / 8 : 0;  // <= n_x, as h <= g
        const size_t span = std::max(h_x, size_t(kHeavySpan * double(n_x)));
        // Costly block j takes a slot in the j-th of h_x equal strata of the
        // span (at least one slot each, as span >= h_x): its first slot,
        // unless that regular spacing has a period that would send the
        // costly blocks to only some of the XCD's 32 CUs (a spacing of
        // exactly 4 slots gave 8 of them all the costly blocks and the launch
        // 1.5x longer); then the slot at the fraction frac(j phi) of the
        // stratum (phi the golden ratio), evenly spread with no period.
        size_t bins[32] = {};
        for (size_t j = 0; j < h_x; ++j) {
            bins[(j * span / h_x) % 32]++;
        }
        bool periodic = false;
        for (size_t c = 0; c < 32; ++c) {
            periodic = periodic || double(bins[c]) > 1.5 * double(h_x) / 32.0 + 2.0;
        }
        for (size_t j = 0; j < h_x; ++j) {
            const size_t lo = j * span / h_x, hi = (j + 1) * span / h_x;
            const double u = double(j) * 0.6180339887498949;
            const size_t r = periodic ? lo + size_t((u - std::floor(u)) * double(hi - lo)) : lo;
            const size_t t = 8 * r + x;
            heavy_slot[t] = 1;
            order[t] = heavy[8 * j + x];
        }
        if (x + 1 < 8) {
            start[x + 1] = start[x] + (n_x - h_x);
        }
    }
    for (size_t t = 0; t < g; ++t) {
        if (!heavy_slot[t]) {
            order[t] = rest[start[t % 8]++];
        }
    }
    // a permutation of the blocks: every element is processed once
    std::vector<char> seen(g, 0);
    for (int32_t b : order) {
        if (b < 0 || size_t(b) >= g || seen[size_t(b)]++) {
            throw std::logic_error("eos_block_order: not a permutation");
        }
    }
    return order;
}

// Recursive coordinate bisection: a range of elements larger than one block
// is cut at the median centroid along its longest extent, into parts holding
// whole numbers of blocks (the first ceil(k / 2) of the range's k blocks),
// so every block but the last of a range is full.  Ties are broken by the
// element index, so the cut is deterministic.  Blocks come out in
// depth-first order, which keeps consecutive blocks near each other.
ForceBlocksU build_force_blocks(const HostMeshU &m, int block_elems)
{
    if (block_elems < 1 || block_elems > kForceBlockElems) {
        throw std::invalid_argument("build_force_blocks: block size out of range");
    }
    const size_t ne = size_t(m.num_elem), nn = size_t(m.num_node);
    const size_t cap = size_t(block_elems);
    std::vector<double> cen[3];
    const std::vector<double> *xyz[3] = {&m.x, &m.y, &m.z};
    for (int a = 0; a < 3; ++a) {
        cen[a].assign(ne, 0.0);
        for (size_t el = 0; el < ne; ++el) {
            double sum = 0.0;
            for (int c = 0; c < 8; ++c) {
                sum += (*xyz[a])[size_t(m.nodelist[size_t(c) * ne + el])];
            }
            cen[a][el] = sum * 0.125;
        }
    }
    std::vector<int32_t> idx(ne);
    for (size_t el = 0; el < ne; ++el) idx[el] = int32_t(el);
    ForceBlocksU b;
    b.elem_start.push_back(0);
    std::vector<std::pair<size_t, size_t>> stack;
    if (ne > 0) stack.emplace_back(0, ne);
    while (!stack.empty()) {
        const auto [lo, hi] = stack.back();
        stack.pop_back();
        const size_t count = hi - lo;
        if (count <= cap) {
            std::sort(idx.begin() + std::ptrdiff_t(lo), idx.begin() + std::ptrdiff_t(hi));
            b.elem_start.push_back(int32_t(hi));
            continue;
        }
        int axis = 0;
        double best = -1.0;
        for (int a = 0; a < 3; ++a) {
            double mn = cen[a][size_t(idx[lo])], mx = mn;
            for (size_t i = lo; i < hi; ++i) {
                mn = std::min(mn, cen[a][size_t(idx[i])]);
                mx = std::max(mx, cen[a][size_t(idx[i])]);
            }
            if (mx - mn > best) {
                best = mx - mn;
                axis = a;
            }
        }
        const size_t blocks = (count + cap - 1) / cap;
        const size_t mid = lo + (blocks + 1) / 2 * cap;
        const std::vector<double> &key = cen[axis];
        std::nth_element(idx.begin() + std::ptrdiff_t(lo), idx.begin() + std::ptrdiff_t(mid),
                         idx.begin() + std::ptrdiff_t(hi), [&](int32_t p, int32_t q) {
                             const double kp = key[size_t(p)], kq = key[size_t(q)];
                             return kp < kq || (kp == kq && p < q);
                         });
        stack.emplace_back(mid, hi);  // popped second: the lower half first
        stack.emplace_back(lo, mid);
    }
    b.num_blocks = int(b.elem_start.size()) - 1;
    // new element numbers: block by block, ascending within a block
    b.elem_new.assign(ne, 0);
    std::vector<int32_t> block_of(ne);
    for (int k = 0; k < b.num_blocks; ++k) {
        for (int32_t i = b.elem_start[size_t(k)]; i < b.elem_start[size_t(k) + 1]; ++i) {
            b.elem_new[size_t(idx[size_t(i)])] = i;
            block_of[size_t(idx[size_t(i)])] = k;
        }
    }
    // interior nodes by block, ascending; a node with no element is a
    // boundary node
    std::vector<std::vector<int32_t>> interior(size_t(b.num_blocks));
    std::vector<char> is_boundary(nn, 1);
    for (size_t n = 0; n < nn; ++n) {
        const int32_t c0 = m.corner_start[n], c1 = m.corner_start[n + 1];
        if (c0 == c1) continue;
        const int k = block_of[size_t(m.corner_list[size_t(c0)]) % ne];
        bool one = true;
        for (int32_t i = c0 + 1; i < c1 && one; ++i) {
            one = block_of[size_t(m.corner_list[size_t(i)]) % ne] == k;
        }
        if (one && interior[size_t(k)].size() < cap) {
            interior[size_t(k)].push_back(int32_t(n));
            is_boundary[n] = 0;
        }
    }
    // new node numbers: the interior nodes block by block, then the boundary
    // nodes by the first new number of their elements (near the blocks whose
    // corner forces they read), ties by their own number
    b.node_new.assign(nn, 0);
    b.node_start.push_back(0);
    b.ncorner_start.push_back(0);
    int32_t next = 0;
    for (int k = 0; k < b.num_blocks; ++k) {
        const int32_t e0 = b.elem_start[size_t(k)];
        for (int32_t n : interior[size_t(k)]) {
            b.node_new[size_t(n)] = next++;
            for (int32_t i = m.corner_start[size_t(n)]; i < m.corner_start[size_t(n) + 1]; ++i) {
                const size_t cl = size_t(m.corner_list[size_t(i)]);
                b.ncorners.push_back(
                    uint16_t(8 * (b.elem_new[cl % ne] - e0) + int32_t(cl / ne)));
            }
            b.ncorner_start.push_back(int32_t(b.ncorners.size()));
        }
        b.node_start.push_back(next);
    }
    std::vector<std::pair<int64_t, int32_t>> bnodes;
    for (size_t n = 0; n < nn; ++n) {
        if (!is_boundary[n]) continue;
        int64_t first = int64_t(ne);
        for (int32_t i = m.corner_start[n]; i < m.corner_start[n + 1]; ++i) {
            first = std::min<int64_t>(first, b.elem_new[size_t(m.corner_list[size_t(i)]) % ne]);
        }
        bnodes.emplace_back(first, int32_t(n));
    }
    std::sort(bnodes.begin(), bnodes.end());
    for (const auto &bn : bnodes) {
        b.node_new[size_t(bn.second)] = next++;
    }
    return b;
}

std::vector<int32_t> boundary_corner_slots(const HostMeshU &r, int32_t first_boundary)
{
    std::vector<int32_t> slot(size_t(r.num_elem) * 8, -1);
    const int32_t base = r.corner_start[size_t(first_boundary)];
    for (int32_t j = base; j < r.corner_start[size_t(r.num_node)]; ++j) {
        slot[size_t(r.corner_list[size_t(j)])] = j - base;
    }
    return slot;
}

HostMeshU renumber_mesh_u(const HostMeshU &m, const ForceBlocksU &b)
{
    const size_t ne = size_t(m.num_elem), nn = size_t(m.num_node);
    HostMeshU r;
    r.num_elem = m.num_elem;
    r.num_node = m.num_node;
    r.origin = b.elem_new[size_t(m.origin)];
    auto by_node = [&](const auto &v) {
        std::remove_cv_t<std::remove_reference_t<decltype(v)>> o(v.size());
        for (size_t n = 0; n < nn; ++n) o[size_t(b.node_new[n])] = v[n];
        return o;
    };
    auto by_elem = [&](const auto &v) {
        std::remove_cv_t<std::remove_reference_t<decltype(v)>> o(v.size());
        for (size_t el = 0; el < ne; ++el) o[size_t(b.elem_new[el])] = v[el];
        return o;
    };
    r.x = by_node(m.x);
    r.y = by_node(m.y);
    r.z = by_node(m.z);
    r.nodal_mass = by_node(m.nodal_mass);
    r.symm = by_node(m.symm);
    r.volo = by_elem(m.volo);
    r.e = by_elem(m.e);
    r.elem_bc = by_elem(m.elem_bc);
    r.nodelist.assign(8 * ne, 0);
    for (size_t c = 0; c < 8; ++c) {
        for (size_t el = 0; el < ne; ++el) {
            r.nodelist[c * ne + size_t(b.elem_new[el])] =
                b.node_new[size_t(m.nodelist[c * ne + el])];
        }
    }
    r.nbr.assign(6 * ne, 0);
    for (size_t f = 0; f < 6; ++f) {
        for (size_t el = 0; el < ne; ++el) {
            r.nbr[f * ne + size_t(b.elem_new[el])] = b.elem_new[size_t(m.nbr[f * ne + el])];
        }
    }
    // each node's corners in their old order, renamed
    r.corner_start.assign(nn + 1, 0);
    for (size_t n = 0; n < nn; ++n) {
        r.corner_start[size_t(b.node_new[n]) + 1] = m.corner_start[n + 1] - m.corner_start[n];
    }
    for (size_t n = 0; n < nn; ++n) r.corner_start[n + 1] += r.corner_start[n];
    r.corner_list.assign(m.corner_list.size(), 0);
    for (size_t n = 0; n < nn; ++n) {
        int32_t out = r.corner_start[size_t(b.node_new[n])];
        for (int32_t i = m.corner_start[n]; i < m.corner_start[n + 1]; ++i) {
            const size_t cl = size_t(m.corner_list[size_t(i)]);
            r.corner_list[size_t(out++)] = int32_t((cl / ne) * ne + size_t(b.elem_new[cl % ne]));
        }
    }
    return r;
}

MeshLists read_lmesh(const char *path)
{
    std::FILE *fp = std::fopen(path, "rb");
    if (!fp) {
        throw std::invalid_argument(std::string("cannot open ") + path);
    }
    std::string text;
    char buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, fp)) > 0) {
        text.append(buf, got);
    }
    std::fclose(fp);
    const char *p = text.c_str();
    const char *end = p + text.size();
    auto next_int = [&](const char *what) -> long {
        char *q = nullptr;
        const long v = std::strtol(p, &q, 10);
        if (q == p || q > end) {
            throw std::invalid_argument(std::string("lmesh: expected ") + what);
        }
        p = q;
        return v;
    };
    auto next_double = [&]() -> double {
        char *q = nullptr;
        const double v = std::strtod(p, &q);
        if (q == p || q > end) {
            throw std::invalid_argument("lmesh: expected a coordinate");
        }
        p = q;
        return v;
    };
    MeshLists l;
    const long ne = next_int("the element count"), nn = next_int("the node count");
    if (ne < 1 || nn < 8 || ne > (1L << 26) || nn > (1L << 27)) {
        throw std::invalid_argument("lmesh: element / node counts out of range");
    }
    l.num_elem = int(ne);
    l.num_node = int(nn);
    l.x.resize(size_t(nn));
    l.y.resize(size_t(nn));
    l.z.resize(size_t(nn));
    for (long i = 0; i < nn; ++i) {
        l.x[size_t(i)] = next_double();
        l.y[size_t(i)] = next_double();
        l.z[size_t(i)] = next_double();
    }
    // indices are range-checked as the long strtol returns, before they are
    // narrowed to int32 (a value such as 2^32 + 5 would otherwise wrap to a
    // valid index)
    auto next_index = [&](const char *what, long limit) -> int32_t {
        const long v = next_int(what);
        if (v < 0 || v >= limit) {
            throw std::invalid_argument(std::string("lmesh: ") + what + " out of range");
        }
        return int32_t(v);
    };
    l.nodelist.resize(size_t(ne) * 8);
    for (auto &v : l.nodelist) v = next_index("a node index", nn);
    l.neighbors.resize(size_t(ne) * 6);
    for (auto &v : l.neighbors) v = next_index("a neighbour index", ne);
    std::vector<int32_t> *sets[4] = {&l.symm_x, &l.symm_y, &l.symm_z, &l.free_surf};
    for (auto *set : sets) {
        const long n = next_int("a node set size");
        if (n < 0 || n > nn) {
            throw std::invalid_argument("lmesh: node set size out of range");
        }
        set->resize(size_t(n));
        for (auto &v : *set) v = next_index("a node set entry", nn);
    }
    return l;
}

HostMeshU build_mesh_u(const MeshLists &l)
{
    const int ne = l.num_elem, nn = l.num_node;
    if (ne > (1 << 26) || nn > (1 << 27)) {
        throw std::invalid_argument("mesh: element / node counts out of range");
    }
    if (ne < 1 || nn < 8 || l.x.size() != size_t(nn) || l.y.size() != size_t(nn) ||
        l.z.size() != size_t(nn) || l.nodelist.size() != size_t(ne) * 8 ||
        l.neighbors.size() != size_t(ne) * 6) {
        throw std::invalid_argument("mesh: array sizes do not match the counts");
    }
    for (int32_t v : l.nodelist) {
        if (v < 0 || v >= nn) throw std::invalid_argument("mesh: node index out of range");
    }
    for (int32_t v : l.neighbors) {
        if (v < 0 || v >= ne) throw std::invalid_argument("mesh: neighbour out of range");
    }
    for (const auto *set : {&l.symm_x, &l.symm_y, &l.symm_z, &l.free_surf}) {
        for (int32_t v : *set) {
            if (v < 0 || v >= nn) throw std::invalid_argument("mesh: node set entry out of range");
        }
    }
    HostMeshU m;
    m.num_elem = ne;
    m.num_node = nn;
    m.x = l.x;
    m.y = l.y;
    m.z = l.z;
    m.nodelist.resize(size_t(ne) * 8);
    m.nbr.resize(size_t(ne) * 6);
    for (int el = 0; el < ne; ++el) {
        for (int c = 0; c < 8; ++c) m.nodelist[size_t(c) * ne + el] = l.nodelist[size_t(el) * 8 + c];
        for (int d = 0; d < 6; ++d) m.nbr[size_t(d) * ne + el] = l.neighbors[size_t(el) * 6 + d];
    }
    // symmetry planes of the nodes
    m.symm.assign(size_t(nn), 0);
    for (int32_t v : l.symm_x) m.symm[size_t(v)] |= 1;
    for (int32_t v : l.symm_y) m.symm[size_t(v)] |= 2;
    for (int32_t v : l.symm_z) m.symm[size_t(v)] |= 4;
    // face masks: a face with its four nodes on the symmetry sets, then (as
    // the reference's reader marks them) a face with its four nodes on the
    // free-surface set, is a symmetry face (lulesh.cu:852-913)
    constexpr int kZM = 0x100, kZP = 0x400, kEM = 0x010, kEP = 0x040, kXM = 0x001, kXP = 0x004;
    m.elem_bc.assign(size_t(ne), 0);
    std::vector<int> surf(static_cast<size_t>(nn));
    for (int pass = 0; pass < 2; ++pass) {
        std::fill(surf.begin(), surf.end(), 0);
        if (pass == 0) {
            for (const auto *set : {&l.symm_x, &l.symm_y, &l.symm_z})
                for (int32_t v : *set) surf[size_t(v)] = 1;
        } else {
            for (int32_t v : l.free_surf) surf[size_t(v)] = 1;
        }
        for (int el = 0; el < ne; ++el) {
            int mask = 0;
            for (int c = 0; c < 8; ++c) mask |= surf[size_t(l.nodelist[size_t(el) * 8 + c])] << c;
            int32_t &bc = m.elem_bc[size_t(el)];
            if ((mask & 0x0f) == 0x0f) bc |= kZM;
            if ((mask & 0xf0) == 0xf0) bc |= kZP;
            if ((mask & 0x33) == 0x33) bc |= kEM;
            if ((mask & 0xcc) == 0xcc) bc |= kEP;
            if ((mask & 0x99) == 0x99) bc |= kXM;
            if ((mask & 0x66) == 0x66) bc |= kXP;
        }
        if (pass == 0) {
            m.origin = 0;
            for (int el = 0; el < ne; ++el) {
                if (m.elem_bc[size_t(el)] == (kXM | kEM | kZM)) {
                    m.origin = el;
                    break;
                }
            }
        }
    }
    // node -> corners in ascending element order
    m.corner_start.assign(size_t(nn) + 1, 0);
    for (int32_t v : l.nodelist) ++m.corner_start[size_t(v) + 1];
    for (int n = 0; n < nn; ++n) m.corner_start[size_t(n) + 1] += m.corner_start[size_t(n)];
    m.corner_list.resize(size_t(ne) * 8);
    {
        std::vector<int32_t> fill(m.corner_start.begin(), m.corner_start.end() - 1);
        for (int el = 0; el < ne; ++el) {
            for (int c = 0; c < 8; ++c) {
                const int32_t n = l.nodelist[size_t(el) * 8 + c];
                m.corner_list[size_t(fill[size_t(n)]++)] = int32_t(c * ne + el);
            }
        }
    }
    // volumes and masses, summed into the nodes in element order
    m.volo.resize(size_t(ne));
    m.nodal_mass.assign(size_t(nn), 0.0);
    for (int el = 0; el < ne; ++el) {
        double xl[8], yl[8], zl[8];
        for (int c = 0; c < 8; ++c) {
            const int32_t n = l.nodelist[size_t(el) * 8 + c];
            xl[c] = l.x[size_t(n)];
            yl[c] = l.y[size_t(n)];
            zl[c] = l.z[size_t(n)];
        }
        const double vol = hex_volume(xl, yl, zl);
        m.volo[size_t(el)] = vol;
        for (int c = 0; c < 8; ++c) {
            m.nodal_mass[size_t(l.nodelist[size_t(el) * 8 + c])] += vol / 8.0;
        }
    }
    m.e.assign(size_t(ne), 0.0);
    m.e[size_t(m.origin)] = 3.948746e+7;
    return m;
}

}  // namespace lulesh
