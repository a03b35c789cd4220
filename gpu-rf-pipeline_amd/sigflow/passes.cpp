// Graph passes of sigflowc: dead-node elimination, FIR fusion, the launch schedule, liveness
// and buffer slots, and the memory plan. Each keeps node ids: a removed node stays in the
// vector with live[id] == false, so dumps before and after a pass name the same values.
//
// Semantics follow the reference's passes (compiler/src/dce.cpp, fuse.cpp, schedule.cpp,
// liveness.cpp), which its golden files pin (compiler/tests/golden/{dce,fuse,schedule,
// liveness}_*.expected):
//   - dce: a node is live if an emit reaches it through inputs;
//   - fusion: FIRs are taken in source order, each joins the first earlier group with the same
//     taps file that has room (kMaxFusedFirs), has no path to or from it, and stays acyclic as
//     a launch graph; groups of one are not groups;
//   - schedule: Kahn's algorithm over launch units (a fusion group is one unit), the ready unit
//     earliest in the source first, a group's members in source order;
//   - slots: linear scan in schedule order; a value takes the lowest free slot, and a slot is
//     free once the launch of its value's last reader is over.

#include "sigflow.hpp"

#include <algorithm>
#include <cstdint>
#include <filesystem>
#include <set>

namespace sigflow {

namespace {

std::vector<std::vector<int>> users_of(const Graph& g)
{
    std::vector<std::vector<int>> users(g.nodes.size());
    for (const Node& n : g.nodes) {
        if (!g.live[static_cast<std::size_t>(n.id)]) {
            continue;
        }
        for (int in : n.inputs) {
            users[static_cast<std::size_t>(in)].push_back(n.id);
        }
    }
    return users;
}

// Reachability along input -> user edges as one bit set of descendants per node, built in
// decreasing id order (every user has a larger id than its inputs): n^2 / 8 bytes, which the
// parser's bound on operations keeps to a few MB.
class Reach {
public:
    Reach(const Graph& g, const std::vector<std::vector<int>>& users)
        : words_((g.nodes.size() + 63) / 64), bits_(g.nodes.size() * words_, 0)
    {
        for (std::size_t v = g.nodes.size(); v-- > 0;) {
            std::uint64_t* dv = row(v);
            for (int u : users[v]) {
                const std::uint64_t* du = row(static_cast<std::size_t>(u));
                for (std::size_t w = 0; w < words_; ++w) {
                    dv[w] |= du[w];
                }
                dv[static_cast<std::size_t>(u) / 64] |= std::uint64_t(1) << (u % 64);
            }
        }
    }
    bool operator()(int from, int to) const
    {
        return (bits_[static_cast<std::size_t>(from) * words_ + static_cast<std::size_t>(to) / 64]
                >> (to % 64)) & 1;
    }

private:
    std::uint64_t* row(std::size_t v) { return bits_.data() + v * words_; }
    std::size_t words_;
    std::vector<std::uint64_t> bits_;
};

// Launch units: unit[i] is the representative of node i's unit (itself outside a group).
// in_degree counts edges into a unit from other units (and self-loops, which only a cycle has);
// next lists a user unit once per such edge.
struct UnitGraph {
    std::vector<int> in_degree;
    std::vector<std::vector<int>> next;
};

UnitGraph unit_graph(const Graph& g, const std::vector<int>& unit)
{
    UnitGraph ug;
    ug.in_degree.assign(g.nodes.size(), 0);
    ug.next.assign(g.nodes.size(), {});
    for (const Node& n : g.nodes) {
        if (!g.live[static_cast<std::size_t>(n.id)]) {
            continue;
        }
        const int to = unit[static_cast<std::size_t>(n.id)];
        for (int in : n.inputs) {
            const int from = unit[static_cast<std::size_t>(in)];
            if (g.live[static_cast<std::size_t>(in)] && (from != to || in == n.id)) {
                ++ug.in_degree[static_cast<std::size_t>(to)];
                ug.next[static_cast<std::size_t>(from)].push_back(to);
            }
        }
    }
    return ug;
}

bool units_acyclic(const Graph& g, const std::vector<int>& unit)
{
    UnitGraph ug = unit_graph(g, unit);
    std::vector<int> ready;
    std::size_t units = 0;
    for (const Node& n : g.nodes) {
        if (g.live[static_cast<std::size_t>(n.id)] && unit[static_cast<std::size_t>(n.id)] == n.id) {
            ++units;
            if (ug.in_degree[static_cast<std::size_t>(n.id)] == 0) {
                ready.push_back(n.id);
            }
        }
    }
    std::size_t done = 0;
    while (!ready.empty()) {
        const int u = ready.back();
        ready.pop_back();
        ++done;
        for (int v : ug.next[static_cast<std::size_t>(u)]) {
            if (--ug.in_degree[static_cast<std::size_t>(v)] == 0) {
                ready.push_back(v);
            }
        }
    }
    return done == units;
}

}  // namespace

void eliminate_dead_nodes(Graph& g)
{
    std::vector<char> reached(g.nodes.size(), 0);
    std::vector<int> todo;
    for (const Node& n : g.nodes) {
        if (n.op == Op::Emit && g.live[static_cast<std::size_t>(n.id)]) {
            todo.push_back(n.id);
        }
    }
    while (!todo.empty()) {
        const int v = todo.back();
        todo.pop_back();
        if (reached[static_cast<std::size_t>(v)] != 0) {
            continue;
        }
        reached[static_cast<std::size_t>(v)] = 1;
        for (int in : g.nodes[static_cast<std::size_t>(v)].inputs) {
            todo.push_back(in);
        }
    }
    for (std::size_t i = 0; i < g.nodes.size(); ++i) {
        g.live[i] = g.live[i] && reached[i] != 0;
    }
}

void fuse_firs(Graph& g)
{
    const std::vector<std::vector<int>> users = users_of(g);
    const Reach reaches(g, users);
    std::vector<int> firs;
    for (const Node& n : g.nodes) {
        if (g.live[static_cast<std::size_t>(n.id)] && n.op == Op::Fir) {
            firs.push_back(n.id);
        }
    }
    std::sort(firs.begin(), firs.end(), [&g](int a, int b) { return g.before(a, b); });

    std::vector<int> unit(g.nodes.size());
    for (std::size_t i = 0; i < unit.size(); ++i) {
        unit[i] = static_cast<int>(i);
    }
    std::vector<std::vector<int>> groups;
    for (int f : firs) {
        const std::string& taps = g.nodes[static_cast<std::size_t>(f)].taps;
        bool placed = false;
        for (std::vector<int>& grp : groups) {
            if (g.nodes[static_cast<std::size_t>(grp.front())].taps != taps
                || static_cast<int>(grp.size()) >= kMaxFusedFirs) {
                continue;
            }
            bool independent = true;
            for (int m : grp) {
                if (reaches(m, f) || reaches(f, m)) {
                    independent = false;
                    break;
                }
            }
            if (!independent) {
                continue;
            }
            unit[static_cast<std::size_t>(f)] = grp.front();
            if (!units_acyclic(g, unit)) {
                unit[static_cast<std::size_t>(f)] = f;
                continue;
            }
            grp.push_back(f);
            placed = true;
            break;
        }
        if (!placed) {
            groups.push_back({f});
        }
    }
    int next = 0;
    for (const std::vector<int>& grp : groups) {
        if (grp.size() < 2) {
            continue;
        }
        for (int m : grp) {
            g.nodes[static_cast<std::size_t>(m)].group = next;
        }
        ++next;
    }
}

Schedule schedule(const Graph& g)
{
    const std::size_t count = g.nodes.size();
    const auto live = [&g](int id) { return static_cast<bool>(g.live[static_cast<std::size_t>(id)]); };

    // The unit of a group is named by its member earliest in the source.
    std::vector<int> unit(count);
    for (std::size_t i = 0; i < count; ++i) {
        unit[i] = static_cast<int>(i);
    }
    std::map<int, int> head;  // group -> representative
    for (const Node& n : g.nodes) {
        if (live(n.id) && n.group >= 0) {
            const auto it = head.find(n.group);
            if (it == head.end() || g.before(n.id, it->second)) {
                head[n.group] = n.id;
            }
        }
    }
    std::vector<std::vector<int>> members(count);
    for (const Node& n : g.nodes) {
        if (!live(n.id)) {
            continue;
        }
        if (n.group >= 0) {
            unit[static_cast<std::size_t>(n.id)] = head.at(n.group);
        }
        members[static_cast<std::size_t>(unit[static_cast<std::size_t>(n.id)])].push_back(n.id);
    }
    UnitGraph ug = unit_graph(g, unit);

    const auto earlier = [&g](int a, int b) { return g.before(a, b); };
    std::set<int, decltype(earlier)> ready(earlier);
    for (const Node& n : g.nodes) {
        if (live(n.id) && unit[static_cast<std::size_t>(n.id)] == n.id
            && ug.in_degree[static_cast<std::size_t>(n.id)] == 0) {
            ready.insert(n.id);
        }
    }
    Schedule s;
    int launch = 0;
    while (!ready.empty()) {
        const int u = *ready.begin();
        ready.erase(ready.begin());
        std::vector<int> ids = members[static_cast<std::size_t>(u)];
        std::sort(ids.begin(), ids.end(), earlier);
        for (int id : ids) {
            s.order.push_back(id);
            s.launch.push_back(launch);
        }
        ++launch;
        for (int v : ug.next[static_cast<std::size_t>(u)]) {
            if (--ug.in_degree[static_cast<std::size_t>(v)] == 0) {
                ready.insert(v);
            }
        }
    }
    for (const Node& n : g.nodes) {
        if (live(n.id) && ug.in_degree[static_cast<std::size_t>(unit[static_cast<std::size_t>(n.id)])] > 0) {
            s.stuck.push_back(n.id);
        }
    }
    return s;
}

Liveness liveness(const Graph& g, const Schedule& s, bool reuse, bool hold_emitted)
{
    const std::size_t count = g.nodes.size();
    std::vector<int> step(count, -1);
    for (std::size_t k = 0; k < s.order.size(); ++k) {
        step[static_cast<std::size_t>(s.order[k])] = static_cast<int>(k);
    }
    std::vector<int> last(step);
    for (int id : s.order) {
        for (int in : g.nodes[static_cast<std::size_t>(id)].inputs) {
            last[static_cast<std::size_t>(in)] =
                std::max(last[static_cast<std::size_t>(in)], step[static_cast<std::size_t>(id)]);
        }
    }
    if (hold_emitted && !s.order.empty()) {
        for (int id : s.order) {
            const Node& n = g.nodes[static_cast<std::size_t>(id)];
            if (n.op == Op::Emit) {
                last[static_cast<std::size_t>(n.inputs[0])] = static_cast<int>(s.order.size()) - 1;
            }
        }
    }
    const auto launch_of = [&s](int k) { return s.launch[static_cast<std::size_t>(k)]; };

    Liveness l;
    l.slot_of.assign(count, -1);
    std::set<int> free_slots;
    std::set<std::pair<int, int>> held;  // (launch of the last reader, slot)
    for (int id : s.order) {
        if (g.nodes[static_cast<std::size_t>(id)].op == Op::Emit) {
            continue;
        }
        const int k = step[static_cast<std::size_t>(id)];
        while (!held.empty() && held.begin()->first < launch_of(k)) {
            free_slots.insert(held.begin()->second);
            held.erase(held.begin());
        }
        int slot = l.slots;
        if (reuse && !free_slots.empty()) {
            slot = *free_slots.begin();
            free_slots.erase(free_slots.begin());
        } else {
            ++l.slots;
        }
        const int end = last[static_cast<std::size_t>(id)];
        held.insert({launch_of(end), slot});
        l.slot_of[static_cast<std::size_t>(id)] = slot;
        l.ranges.push_back(LiveRange{id, k, end, slot});
    }
    return l;
}

std::vector<long long> value_bytes(const Graph& g, const std::string& input_dir)
{
    std::vector<long long> samples(g.nodes.size(), -1);
    std::vector<long long> bytes(g.nodes.size(), -1);
    for (const Node& n : g.nodes) {
        const auto in = [&](std::size_t i) { return samples[static_cast<std::size_t>(n.inputs[i])]; };
        long long count = -1;
        switch (n.op) {
        case Op::Capture: {
            std::error_code ec;
            const auto size = std::filesystem::file_size(std::filesystem::path(input_dir) / n.path, ec);
            if (!ec && size % 8 == 0) {
                count = static_cast<long long>(size / 8);
            }
            break;
        }
        case Op::Fir:
            count = in(0);
            break;
        case Op::Magnitude:
            count = in(0);
            if (n.spectrum_of >= 0 && count >= 0) {
                count = (count + n.type.n - 1) / n.type.n * n.type.n;
            }
            break;
        case Op::Fft:
        case Op::Channelize:
            count = in(0) < 0 ? -1 : (in(0) + n.n - 1) / n.n * n.n;
            break;
        case Op::Xcorr:
        case Op::GccPhat:
            count = in(0) < 0 || in(1) < 0 ? -1 : in(0) + in(1) - 1;
            break;
        case Op::Emit:
            bytes[static_cast<std::size_t>(n.id)] = 0;
            continue;
        }
        samples[static_cast<std::size_t>(n.id)] = count;
        bytes[static_cast<std::size_t>(n.id)] = count < 0 ? -1 : count * (n.type.real ? 4 : 8);
    }
    return bytes;
}

MemoryPlan plan_memory(const Liveness& l, const std::vector<long long>& bytes)
{
    MemoryPlan plan;
    int steps = 0;
    for (const LiveRange& r : l.ranges) {
        if (bytes[static_cast<std::size_t>(r.node)] < 0) {
            return plan;
        }
        steps = std::max(steps, r.last + 1);
    }
    std::vector<long long> slot_size(static_cast<std::size_t>(l.slots), 0);
    std::vector<long long> live_at(static_cast<std::size_t>(steps), 0);
    plan.no_reuse = 0;
    for (const LiveRange& r : l.ranges) {
        const long long b = bytes[static_cast<std::size_t>(r.node)];
        plan.no_reuse += b;
        long long& s = slot_size[static_cast<std::size_t>(r.slot)];
        s = std::max(s, b);
        for (int k = r.def; k <= r.last; ++k) {
            live_at[static_cast<std::size_t>(k)] += b;
        }
    }
    plan.reuse = 0;
    for (long long s : slot_size) {
        plan.reuse += s;
    }
    plan.live_peak = live_at.empty() ? 0 : *std::max_element(live_at.begin(), live_at.end());
    return plan;
}

// An fft whose only live reader is a magnitude, and which nothing emits, folds into that
// magnitude: one spectrum-magnitude call reads the fft's input and writes |X|, and the spectrum
// itself is never stored. Every length: the library's last pass of each plan kind (the
// one-workgroup transform, the four-step row pass, the DFT) stores the magnitude.
void fuse_spectrum(Graph& g)
{
    const std::vector<std::vector<int>> users = users_of(g);
    for (Node& m : g.nodes) {
        if (!g.live[static_cast<std::size_t>(m.id)] || m.op != Op::Magnitude) {
            continue;
        }
        const int f = m.inputs[0];
        const Node& fft = g.nodes[static_cast<std::size_t>(f)];
        if (fft.op != Op::Fft || users[static_cast<std::size_t>(f)].size() != 1) {
            continue;
        }
        m.spectrum_of = f;
        m.inputs = fft.inputs;
        g.live[static_cast<std::size_t>(f)] = false;
    }
}

}  // namespace sigflow
