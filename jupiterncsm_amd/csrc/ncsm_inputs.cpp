// ncsm_inputs.cpp - synthetic M-scheme basis, index lists, matrix-element
// blocks and evaluation order in the reference's formats.
//
// Conventions (the reference consumes them; the external producers are not
// part of the reference tree):
//  * orbitals per species in Antoine order: e, l, j ascending, m descending
//    (src/Neptune/bases/shells.c:72-131, sp_states.c:7-35); global sp index
//    of proton orbital s is 2s, of neutron orbital s is 2s+1
//    (src/Mercury/connection_list/connection_list.c:352-369);
//  * a k-particle state of one species is an ascending orbital tuple; the
//    per-energy state lists ("basis_energy_<E>") keep equal-M states
//    contiguous (M descending, then lexicographic), which the connection
//    enumeration of connection_list.c:425-446 relies on;
//  * a connection (annihilated C -> created A) of a matrix-element block with
//    (depth d, dE, dM) has C at energy d and A at energy d+dE with
//    M(A) = M(C)+dM; elements are annihilated-major, created-minor
//    (connection_list.c:409-441); neutron-proton blocks are
//    neutron-connection + Dn * proton-connection (connection_list.c:138-160,
//    matrix_vector_multiplication.c:186-189);
//  * the k-body operator of a connection is a+_{A1}..a+_{Ak} a_{Ck}..a_{C1};
//    index-list entries carry its fermionic phase in bit 31
//    (matrix_vector_multiplication.c:7-18);
//  * vector element (p, n) of a basis block sits at n + Nn*p
//    (matrix_vector_multiplication.c:58-63).
#include "ncsm_inputs.h"

#include <algorithm>
#include <array>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <sys/mman.h>
#include <sys/stat.h>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "ncsm_gpu.h"

namespace {

constexpr int kMaxParticles = 4;
using Tuple = std::array<int16_t, kMaxParticles>;

struct Orbital {
	int e, l, j2, m2;
};

uint64_t tuple_key(const Tuple &t, int k)
{
	uint64_t key = 0;
	for (int i = 0; i < k; i++)
		key |= uint64_t(uint16_t(t[i] + 1)) << (16 * i);
	return key;
}

// Fixed-energy list of k-particle states of one species, M descending then
// lexicographic, so equal-M states are contiguous.
struct StateSet {
	std::vector<Tuple> states;
	std::vector<int> m2;
	std::unordered_map<uint64_t, int> index;
	std::map<int, std::pair<int, int>> mgroup;  // 2M -> (first, count)
};

// All Slater determinants of one species with fixed (E, 2M), lexicographic.
struct SubBasis {
	int E = 0, m2 = 0;
	std::vector<Tuple> dets;
	std::unordered_map<uint64_t, int> index;
};

struct Species {
	int k = 0;
	int emin = 0, emax = 0;
	std::vector<SubBasis> subs;
	std::map<std::pair<int, int>, int> sub_of;  // (E, 2M) -> sub index
	std::map<std::pair<int, int>, StateSet> kstates;  // (k', E) -> states
};

// Offsets of annihilated states into the connection enumeration of one
// (k', d, dE, dM) species part.
struct ConnTable {
	std::vector<int64_t> offset;
	int64_t total = 0;
};

struct ListKey {
	int species, k, sub_in, sub_out;
	bool operator<(const ListKey &o) const
	{
		return std::tie(species, k, sub_in, sub_out) <
		       std::tie(o.species, o.k, o.sub_in, o.sub_out);
	}
};

struct ListData {
	int species, k, E_bra, m2_bra, E_ket, m2_ket, depth;
	std::vector<int32_t> entries;  // {in,out,matrix} triples
};

struct VKey {
	int kp, kn;            // particles of the proton / neutron part
	int dp, dEp, dMp;      // proton depth and differences
	int dn, dEn, dMn;      // neutron depth and differences
	bool operator<(const VKey &o) const
	{
		return std::tie(kp, kn, dp, dEp, dMp, dn, dEn, dMn) <
		       std::tie(o.kp, o.kn, o.dp, o.dEp, o.dMp, o.dn, o.dEn,
				o.dMn);
	}
};

uint64_t mix64(uint64_t z)
{
	z += 0x9e3779b97f4a7c15ULL;
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
	return z ^ (z >> 31);
}

// Standard normal variate keyed by (seed, rank, {bra, ket}) through
// splitmix64 and the Box-Muller transform; symmetric in bra <-> ket.
double keyed_gaussian(uint64_t seed, int rank, const int32_t *bra,
		      const int32_t *ket, int n)
{
	const int32_t *lo = bra, *hi = ket;
	if (std::lexicographical_compare(ket, ket + n, bra, bra + n))
		std::swap(lo, hi);
	uint64_t h = mix64(seed ^ (uint64_t(rank) * 0x100000001b3ULL));
	for (int i = 0; i < n; i++)
		h = mix64(h ^ uint64_t(uint32_t(lo[i])));
	h = mix64(h ^ 0x5bd1e995ULL);
	for (int i = 0; i < n; i++)
		h = mix64(h ^ uint64_t(uint32_t(hi[i])));
	const double u1 = double((h >> 11) + 1) * 0x1.0p-53;
	const double u2 = double(mix64(h) >> 11) * 0x1.0p-53;
	return std::sqrt(-2.0 * std::log(u1)) *
	       std::cos(6.283185307179586476925286766559 * u2);
}

}  // namespace

namespace {
// allocator whose resize() leaves new elements uninitialized: the element
// array is written in full afterwards, so zeroing gigabytes first would
// only add a pass
template <typename T> struct NoInitAlloc : std::allocator<T> {
	template <typename U> struct rebind {
		using other = NoInitAlloc<U>;
	};
	NoInitAlloc() = default;
	template <typename U> NoInitAlloc(const NoInitAlloc<U> &) noexcept {}
	template <typename U> void construct(U *q) noexcept { ::new ((void *)q) U; }
	template <typename U, typename... A> void construct(U *q, A &&...a)
	{
		::new ((void *)q) U(std::forward<A>(a)...);
	}
};
}  // namespace

struct ncsm_inputs {
	ncsm_config_t cfg{};
	std::vector<Orbital> orbs;
	int emax_sp = 0;
	int e_total_max = 0;
	Species sp[2];  // 0 = proton, 1 = neutron

	// basis blocks
	std::vector<int64_t> blocks;  // Ep, 2Mp, En, 2Mn, Np, Nn
	std::vector<int> block_sub_p, block_sub_n;
	int64_t dimension = 0;

	std::map<std::tuple<int, int, int, int, int>, ConnTable> conn_tables;

	std::vector<ListData> lists;
	std::vector<int64_t> list_offsets;
	std::vector<int32_t> list_entries;
	std::vector<int32_t> list_settings;

	std::vector<VKey> vkeys;
	std::vector<int64_t> v_offsets;
	std::vector<int64_t> v_shapes;  // Dn, Dp, force
	std::vector<double, NoInitAlloc<double>> v_elements;

	std::vector<ncsm_instruction_t> instr;
	std::vector<int64_t> instr_mults;

	std::vector<int32_t> orbital_table;
};

namespace {

int min_energy(const std::vector<Orbital> &orbs, int k)
{
	int e = 0;
	for (int i = 0; i < k && i < (int)orbs.size(); i++)
		e += orbs[i].e;
	return e;
}

void enumerate_tuples(const std::vector<Orbital> &orbs, int k, int ebound,
		      int start, int depth, int esum, Tuple &cur,
		      std::vector<Tuple> &out)
{
	if (depth == k) {
		out.push_back(cur);
		return;
	}
	const int remaining = k - depth;
	for (int o = start; o < (int)orbs.size(); o++) {
		if (esum + orbs[o].e * remaining > ebound)
			break;
		cur[depth] = (int16_t)o;
		enumerate_tuples(orbs, k, ebound, o + 1, depth + 1,
				 esum + orbs[o].e, cur, out);
	}
}

int tuple_energy(const std::vector<Orbital> &orbs, const Tuple &t, int k)
{
	int e = 0;
	for (int i = 0; i < k; i++)
		e += orbs[t[i]].e;
	return e;
}

int tuple_m2(const std::vector<Orbital> &orbs, const Tuple &t, int k)
{
	int m = 0;
	for (int i = 0; i < k; i++)
		m += orbs[t[i]].m2;
	return m;
}

void build_species(ncsm_inputs &in, Species &s, int k, int emin, int emax,
		   int max_body)
{
	s.k = k;
	s.emin = emin;
	s.emax = emax;
	std::vector<Tuple> all;
	Tuple cur{};
	cur.fill(-1);
	enumerate_tuples(in.orbs, k, emax, 0, 0, 0, cur, all);
	// lexicographic order is the enumeration order
	for (const Tuple &t : all) {
		const int E = tuple_energy(in.orbs, t, k);
		if (E < emin)
			continue;
		const int m2 = tuple_m2(in.orbs, t, k);
		auto it = s.sub_of.find({E, m2});
		int idx;
		if (it == s.sub_of.end()) {
			idx = (int)s.subs.size();
			s.sub_of[{E, m2}] = idx;
			s.subs.emplace_back();
			s.subs.back().E = E;
			s.subs.back().m2 = m2;
		} else
			idx = it->second;
		SubBasis &sb = s.subs[idx];
		sb.index[tuple_key(t, k)] = (int)sb.dets.size();
		sb.dets.push_back(t);
	}
	// k'-particle state sets for every operator rank used on this species
	for (int kk = 1; kk <= std::min(k, max_body); kk++) {
		const int bound = emax - min_energy(in.orbs, k - kk);
		std::vector<Tuple> states;
		Tuple c{};
		c.fill(-1);
		enumerate_tuples(in.orbs, kk, bound, 0, 0, 0, c, states);
		std::map<int, std::vector<Tuple>> by_energy;
		for (const Tuple &t : states)
			by_energy[tuple_energy(in.orbs, t, kk)].push_back(t);
		for (auto &ev : by_energy) {
			std::vector<Tuple> &v = ev.second;
			std::stable_sort(v.begin(), v.end(),
					 [&](const Tuple &a, const Tuple &b) {
						 return tuple_m2(in.orbs, a, kk) >
							tuple_m2(in.orbs, b, kk);
					 });
			StateSet &ss = s.kstates[{kk, ev.first}];
			ss.states = v;
			for (size_t i = 0; i < v.size(); i++) {
				const int m2 = tuple_m2(in.orbs, v[i], kk);
				ss.m2.push_back(m2);
				ss.index[tuple_key(v[i], kk)] = (int)i;
				auto g = ss.mgroup.find(m2);
				if (g == ss.mgroup.end())
					ss.mgroup[m2] = {(int)i, 1};
				else
					g->second.second++;
			}
		}
	}
}

const StateSet *find_states(const Species &s, int kk, int E)
{
	auto it = s.kstates.find({kk, E});
	return it == s.kstates.end() ? nullptr : &it->second;
}

const ConnTable &conn_table(ncsm_inputs &in, int species, int kk, int d,
			    int dE, int dM)
{
	auto key = std::make_tuple(species, kk, d, dE, dM);
	auto it = in.conn_tables.find(key);
	if (it != in.conn_tables.end())
		return it->second;
	ConnTable t;
	const StateSet *ann = find_states(in.sp[species], kk, d);
	const StateSet *cre = find_states(in.sp[species], kk, d + dE);
	if (ann) {
		t.offset.resize(ann->states.size());
		for (size_t i = 0; i < ann->states.size(); i++) {
			t.offset[i] = t.total;
			if (!cre)
				continue;
			auto g = cre->mgroup.find(ann->m2[i] + dM);
			if (g != cre->mgroup.end())
				t.total += g->second.second;
		}
	}
	return in.conn_tables.emplace(key, std::move(t)).first->second;
}

// Phase of a+_{A1}..a+_{Ak} a_{Ck}..a_{C1} acting on the ascending
// determinant det (annihilate C1 first, then create Ak first).
int operator_phase(const Tuple &det, int k, const Tuple &C, const Tuple &A,
		   int kk)
{
	int occ[2 * kMaxParticles];
	int n = k;
	for (int i = 0; i < k; i++)
		occ[i] = det[i];
	int sign = 1;
	for (int c = 0; c < kk; c++) {
		int pos = 0;
		while (occ[pos] != C[c])
			pos++;
		if (pos & 1)
			sign = -sign;
		for (int i = pos; i < n - 1; i++)
			occ[i] = occ[i + 1];
		n--;
	}
	for (int a = kk - 1; a >= 0; a--) {
		int pos = 0;
		while (pos < n && occ[pos] < A[a])
			pos++;
		if (pos & 1)
			sign = -sign;
		for (int i = n; i > pos; i--)
			occ[i] = occ[i - 1];
		occ[pos] = A[a];
		n++;
	}
	return sign;
}

// Index-list entries of a k'-body operator of one species between two
// sub-bases, split by depth (energy of the annihilated k'-tuple).
std::map<int, std::vector<int32_t>> build_list_group(ncsm_inputs &in,
						     const ListKey &key)
{
	std::map<int, std::vector<int32_t>> out;
	const Species &s = in.sp[key.species];
	const SubBasis &sin = s.subs[key.sub_in];
	const SubBasis &sout = s.subs[key.sub_out];
	const int k = s.k, kk = key.k;
	const int dE = sout.E - sin.E, dM = sout.m2 - sin.m2;
	// all k'-subsets of k positions, ascending
	std::vector<std::array<int, kMaxParticles>> subsets;
	{
		std::array<int, kMaxParticles> idx{};
		for (int i = 0; i < kk; i++)
			idx[i] = i;
		while (true) {
			subsets.push_back(idx);
			int i = kk - 1;
			while (i >= 0 && idx[i] == k - kk + i)
				i--;
			if (i < 0)
				break;
			idx[i]++;
			for (int j = i + 1; j < kk; j++)
				idx[j] = idx[j - 1] + 1;
		}
	}
	for (size_t din = 0; din < sin.dets.size(); din++) {
		const Tuple &det = sin.dets[din];
		for (const auto &sub : subsets) {
			Tuple C{}, R{};
			C.fill(-1);
			R.fill(-1);
			int nr = 0;
			for (int i = 0, c = 0; i < k; i++) {
				if (c < kk && sub[c] == i)
					C[c++] = det[i];
				else
					R[nr++] = det[i];
			}
			const int d = tuple_energy(in.orbs, C, kk);
			const int mC = tuple_m2(in.orbs, C, kk);
			const StateSet *ann = find_states(s, kk, d);
			const StateSet *cre = find_states(s, kk, d + dE);
			if (!ann || !cre)
				continue;
			auto g = cre->mgroup.find(mC + dM);
			if (g == cre->mgroup.end())
				continue;
			const ConnTable &ct =
				conn_table(in, key.species, kk, d, dE, dM);
			const int ann_pos = ann->index.at(tuple_key(C, kk));
			for (int a = 0; a < g->second.second; a++) {
				const Tuple &A =
					cre->states[g->second.first + a];
				bool clash = false;
				for (int i = 0; i < kk && !clash; i++)
					for (int j = 0; j < nr; j++)
						if (A[i] == R[j])
							clash = true;
				if (clash)
					continue;
				Tuple O{};
				O.fill(-1);
				int no = 0;
				for (int i = 0; i < nr; i++)
					O[no++] = R[i];
				for (int i = 0; i < kk; i++)
					O[no++] = A[i];
				std::sort(O.begin(), O.begin() + k);
				auto oi = sout.index.find(tuple_key(O, k));
				if (oi == sout.index.end())
					continue;
				const int sign = operator_phase(det, k, C, A,
								kk);
				const int64_t m = ct.offset[ann_pos] + a;
				uint32_t mi = (uint32_t)m;
				if (sign < 0)
					mi |= 0x80000000u;
				std::vector<int32_t> &v = out[d];
				v.push_back((int32_t)din);
				v.push_back((int32_t)oi->second);
				v.push_back((int32_t)mi);
			}
		}
	}
	return out;
}

// global orbital tuple of a connection side: protons first (2s), then
// neutrons (2s+1)
int fill_global(int32_t *dst, const Tuple *p, int kp, const Tuple *n, int kn)
{
	int c = 0;
	for (int i = 0; i < kp; i++)
		dst[c++] = 2 * (*p)[i];
	for (int i = 0; i < kn; i++)
		dst[c++] = 2 * (*n)[i] + 1;
	return c;
}

struct ConnList {
	std::vector<std::pair<int, int>> pairs;  // (ann pos, cre pos)
	const StateSet *ann = nullptr, *cre = nullptr;
};

ConnList conn_list(ncsm_inputs &in, int species, int kk, int d, int dE,
		   int dM)
{
	ConnList cl;
	if (kk == 0) {
		cl.pairs.push_back({0, 0});
		return cl;
	}
	cl.ann = find_states(in.sp[species], kk, d);
	cl.cre = find_states(in.sp[species], kk, d + dE);
	if (!cl.ann || !cl.cre)
		return cl;
	for (size_t i = 0; i < cl.ann->states.size(); i++) {
		auto g = cl.cre->mgroup.find(cl.ann->m2[i] + dM);
		if (g == cl.cre->mgroup.end())
			continue;
		for (int a = 0; a < g->second.second; a++)
			cl.pairs.push_back({(int)i, g->second.first + a});
	}
	return cl;
}

bool make_dir(const std::string &path)
{
	if (mkdir(path.c_str(), 0755) == 0 || errno == EEXIST)
		return true;
	fprintf(stderr, "ncsm_inputs: cannot create %s: %s\n", path.c_str(),
		strerror(errno));
	return false;
}

const char *type_string(int kk, char c)
{
	static const char *p[] = {"", "p", "pp", "ppp", "pppp"};
	static const char *n[] = {"", "n", "nn", "nnn", "nnnn"};
	return c == 'p' ? p[kk] : n[kk];
}

}  // namespace

extern "C" {

namespace {

// stage times on stderr when NCSM_VERBOSE is set
struct StageClock {
	bool on = getenv("NCSM_VERBOSE") != nullptr;
	std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
	void mark(const char *what)
	{
		const auto now = std::chrono::steady_clock::now();
		if (on)
			fprintf(stderr, "[ncsm] %s %.3f s\n", what,
				std::chrono::duration<double>(now - t).count());
		t = now;
	}
};

ncsm::Packed pack(const Tuple &t)
{
	ncsm::Packed p = 0;
	for (int i = 0; i < kMaxParticles; i++)
		p |= ncsm::Packed(uint16_t(t[i])) << (16 * i);
	return p;
}

// Every list group's entries on the GPU (the work of build_list_group):
// the host makes one item per (group, input determinant, annihilated
// subset) in build_list_group's loop order, the device returns each item's
// entries in candidate order, and they are appended to the item's depth
// list, which gives the host generator's lists exactly.
bool gpu_list_groups(ncsm_inputs &in, const std::vector<ListKey> &groups,
		     std::vector<std::map<int, std::vector<int32_t>>> &group_lists)
{
	// state sets and sub-basis keys of both species, flattened
	StageClock clk;
	std::map<const StateSet *, int64_t> state_off;
	std::vector<ncsm::Packed> states;
	std::map<const SubBasis *, int64_t> key_off;
	std::vector<uint64_t> keys;
	for (int sp = 0; sp < 2; sp++) {
		for (auto &kv : in.sp[sp].kstates) {
			state_off[&kv.second] = (int64_t)states.size();
			for (const Tuple &t : kv.second.states)
				states.push_back(pack(t));
		}
		for (const SubBasis &sb : in.sp[sp].subs) {
			key_off[&sb] = (int64_t)keys.size();
			for (const Tuple &t : sb.dets)
				keys.push_back(ncsm::key_be(t.data(), in.sp[sp].k));
		}
	}
	std::vector<std::vector<ncsm::GenItem>> per(groups.size());
	std::vector<std::vector<int>> depth(groups.size());
#pragma omp parallel for schedule(dynamic, 1)
	for (long gi = 0; gi < (long)groups.size(); gi++) {
		const ListKey &key = groups[gi];
		const Species &s = in.sp[key.species];
		const SubBasis &sin = s.subs[key.sub_in];
		const SubBasis &sout = s.subs[key.sub_out];
		const int k = s.k, kk = key.k;
		const int dE = sout.E - sin.E, dM = sout.m2 - sin.m2;
		for (size_t din = 0; din < sin.dets.size(); din++) {
			const Tuple &det = sin.dets[din];
			// k'-subsets of the k positions in ascending order, as
			// build_list_group enumerates them
			std::array<int, kMaxParticles> idx{};
			for (int i = 0; i < kk; i++)
				idx[i] = i;
			while (true) {
				Tuple C{};
				C.fill(-1);
				uint8_t mask = 0;
				for (int c = 0; c < kk; c++) {
					C[c] = det[idx[c]];
					mask |= uint8_t(1u << idx[c]);
				}
				const int d = tuple_energy(in.orbs, C, kk);
				const int mC = tuple_m2(in.orbs, C, kk);
				const StateSet *ann = find_states(s, kk, d);
				const StateSet *cre = find_states(s, kk, d + dE);
				if (ann && cre) {
					auto g = cre->mgroup.find(mC + dM);
					if (g != cre->mgroup.end()) {
						const auto ct = in.conn_tables.find(
							std::make_tuple(key.species, kk, d, dE, dM));
						const int ann_pos = ann->index.at(tuple_key(C, kk));
						ncsm::GenItem it{};
						it.det = pack(det);
						it.cre_first = state_off.at(cre) + g->second.first;
						it.m_base = ct->second.offset[ann_pos];
						it.out_first = key_off.at(&sout);
						it.out_count = (int32_t)sout.dets.size();
						it.cre_count = g->second.second;
						it.din = (int32_t)din;
						it.sub = mask;
						it.k = (uint8_t)k;
						it.kk = (uint8_t)kk;
						per[gi].push_back(it);
						depth[gi].push_back(d);
					}
				}
				int i = kk - 1;
				while (i >= 0 && idx[i] == k - kk + i)
					i--;
				if (i < 0)
					break;
				idx[i]++;
				for (int j = i + 1; j < kk; j++)
					idx[j] = idx[j - 1] + 1;
			}
		}
	}
	std::vector<ncsm::GenItem> items;
	for (auto &v : per)
		items.insert(items.end(), v.begin(), v.end());
	clk.mark("list work items (host)");
	std::vector<int32_t> entries, counts;
	std::string err;
	const int rc = ncsm::gpu_list_entries(items, states, keys, entries, counts, err);
	clk.mark("list entries (device)");
	if (rc) {
		fprintf(stderr, "ncsm_inputs_build_gpu: index lists: %s\n", err.c_str());
		return false;
	}
	// every group's first item and first entry, then the groups split in
	// parallel
	std::vector<size_t> item0(groups.size() + 1, 0), entry0(groups.size() + 1, 0);
	for (size_t gi = 0; gi < groups.size(); gi++) {
		item0[gi + 1] = item0[gi] + per[gi].size();
		size_t n = 0;
		for (size_t j = item0[gi]; j < item0[gi + 1]; j++)
			n += counts[j];
		entry0[gi + 1] = entry0[gi] + n;
	}
#pragma omp parallel for schedule(dynamic, 1)
	for (long gi = 0; gi < (long)groups.size(); gi++) {
		size_t e = entry0[gi];
		for (size_t j = 0; j < per[gi].size(); j++) {
			const int32_t n = counts[item0[gi] + j];
			if (!n)
				continue;
			std::vector<int32_t> &v = group_lists[gi][depth[gi][j]];
			v.insert(v.end(), entries.begin() + 3 * e, entries.begin() + 3 * (e + n));
			e += n;
		}
	}
	clk.mark("list split by depth (host)");
	return true;
}

ncsm_inputs_t *build_inputs(const ncsm_config_t *config, bool gpu);

}  // namespace

ncsm_inputs_t *ncsm_inputs_build(const ncsm_config_t *config)
{
	return build_inputs(config, false);
}

ncsm_inputs_t *ncsm_inputs_build_gpu(const ncsm_config_t *config)
{
	return build_inputs(config, true);
}

namespace {

ncsm_inputs_t *build_inputs(const ncsm_config_t *config, bool gpu)
{
	if (!config)
		return nullptr;
	const ncsm_config_t &c = *config;
	if (c.num_protons < 1 || c.num_neutrons < 1 ||
	    c.num_protons > kMaxParticles || c.num_neutrons > kMaxParticles ||
	    c.nmax < 0 || c.max_body < 1 || c.max_body > 3) {
		fprintf(stderr, "ncsm_inputs_build: unsupported configuration "
				"(Z=%d N=%d Nmax=%d max_body=%d)\n",
			c.num_protons, c.num_neutrons, c.nmax, c.max_body);
		return nullptr;
	}
	StageClock clk;
	ncsm_inputs *in = new ncsm_inputs;
	in->cfg = c;
	// single-particle basis large enough for any orbital the Nmax
	// truncation allows; trimmed below once the minimum filling is known
	const int emax_guess = c.nmax + 8;
	std::vector<Orbital> orbs;
	for (int e = 0; e <= emax_guess; e++)
		for (int l = e & 1; l <= e; l += 2)
			for (int j2 = std::abs(2 * l - 1); j2 <= 2 * l + 1;
			     j2 += 2)
				for (int m2 = j2; m2 >= -j2; m2 -= 2)
					orbs.push_back({e, l, j2, m2});
	const int minp = min_energy(orbs, c.num_protons);
	const int minn = min_energy(orbs, c.num_neutrons);
	in->e_total_max = minp + minn + c.nmax;
	// largest single-particle energy: the rest of the nucleus at its
	// lowest filling
	const int e_rest = std::min(min_energy(orbs, c.num_protons - 1) + minn,
				    minp + min_energy(orbs, c.num_neutrons - 1));
	in->emax_sp = in->e_total_max - e_rest;
	for (const Orbital &o : orbs)
		if (o.e <= in->emax_sp)
			in->orbs.push_back(o);
	for (const Orbital &o : in->orbs) {
		in->orbital_table.push_back(o.e);
		in->orbital_table.push_back(o.l);
		in->orbital_table.push_back(o.j2);
		in->orbital_table.push_back(o.m2);
	}
	build_species(*in, in->sp[0], c.num_protons, minp,
		      in->e_total_max - minn, c.max_body);
	build_species(*in, in->sp[1], c.num_neutrons, minn,
		      in->e_total_max - minp, c.max_body);

	// basis blocks: total 2M fixed, parity of the lowest filling,
	// Ep + En <= e_total_max; ordered by Ep, En, then 2Mp descending
	const int parity = (minp + minn) & 1;
	struct B {
		int Ep, mp, En, mn, sp, sn;
	};
	std::vector<B> bl;
	for (size_t i = 0; i < in->sp[0].subs.size(); i++) {
		const SubBasis &p = in->sp[0].subs[i];
		for (size_t j = 0; j < in->sp[1].subs.size(); j++) {
			const SubBasis &n = in->sp[1].subs[j];
			if (p.E + n.E > in->e_total_max)
				continue;
			if (((p.E + n.E) & 1) != parity)
				continue;
			if (p.m2 + n.m2 != c.twice_m)
				continue;
			bl.push_back({p.E, p.m2, n.E, n.m2, (int)i, (int)j});
		}
	}
	std::sort(bl.begin(), bl.end(), [](const B &a, const B &b) {
		return std::make_tuple(a.Ep, a.En, -a.mp) <
		       std::make_tuple(b.Ep, b.En, -b.mp);
	});
	for (const B &b : bl) {
		const int64_t np = in->sp[0].subs[b.sp].dets.size();
		const int64_t nn = in->sp[1].subs[b.sn].dets.size();
		in->blocks.insert(in->blocks.end(),
				  {b.Ep, b.mp, b.En, b.mn, np, nn});
		in->block_sub_p.push_back(b.sp);
		in->block_sub_n.push_back(b.sn);
		in->dimension += np * nn;
	}
	const int nb = (int)bl.size();

	// pass 1: list groups needed by the upper triangle of block pairs
	struct Req {
		int type, rank, bi, bo, kn, kp;
	};
	std::vector<Req> reqs;
	std::map<ListKey, int> group_of;
	std::vector<ListKey> groups;
	auto want = [&](int species, int kk, int si, int so) {
		ListKey key{species, kk, si, so};
		if (!group_of.count(key)) {
			group_of[key] = (int)groups.size();
			groups.push_back(key);
		}
	};
	const int kp_all = c.num_protons, kn_all = c.num_neutrons;
	for (int bi = 0; bi < nb; bi++) {
		for (int bo = bi; bo < nb; bo++) {
			const int pin = in->block_sub_p[bi],
				  pout = in->block_sub_p[bo];
			const int nin = in->block_sub_n[bi],
				  nout = in->block_sub_n[bo];
			if (pin == pout) {
				for (int kk = 1;
				     kk <= std::min(kn_all, c.max_body); kk++) {
					if (kk == 1 &&
					    (bi != bo || !c.include_one_body))
						continue;
					want(1, kk, nin, nout);
					reqs.push_back({1, kk, bi, bo, kk, 0});
				}
			}
			if (nin == nout) {
				for (int kk = 1;
				     kk <= std::min(kp_all, c.max_body); kk++) {
					if (kk == 1 &&
					    (bi != bo || !c.include_one_body))
						continue;
					want(0, kk, pin, pout);
					reqs.push_back({2, kk, bi, bo, 0, kk});
				}
			}
			const int np_ranks[3][2] = {{1, 1}, {2, 1}, {1, 2}};
			for (int r = 0; r < 3; r++) {
				const int kn = np_ranks[r][0],
					  kp = np_ranks[r][1];
				if (kn + kp > c.max_body || kn > kn_all ||
				    kp > kp_all)
					continue;
				want(1, kn, nin, nout);
				want(0, kp, pin, pout);
				reqs.push_back({3, kn + kp, bi, bo, kn, kp});
			}
		}
	}
	// connection tables are created lazily; create every one a list group
	// may need up front so the parallel pass only reads the map
	for (const ListKey &g : groups) {
		const Species &s = in->sp[g.species];
		const int dE = s.subs[g.sub_out].E - s.subs[g.sub_in].E;
		const int dM = s.subs[g.sub_out].m2 - s.subs[g.sub_in].m2;
		for (auto &kv : s.kstates)
			if (kv.first.first == g.k)
				conn_table(*in, g.species, g.k, kv.first.second,
					   dE, dM);
	}
	clk.mark("bases, blocks, list groups, connection tables");
	// pass 2: build every list group
	std::vector<std::map<int, std::vector<int32_t>>> group_lists(
		groups.size());
	if (gpu) {
		if (!gpu_list_groups(*in, groups, group_lists)) {
			delete in;
			return nullptr;
		}
	} else {
#pragma omp parallel for schedule(dynamic, 1)
		for (long gi = 0; gi < (long)groups.size(); gi++)
			group_lists[gi] = build_list_group(*in, groups[gi]);
	}

	clk.mark("index lists");
	// pass 3: ids in order of first use, instructions
	std::map<std::pair<int, int>, int> list_id;  // (group, depth) -> id
	std::map<VKey, int> v_id;
	auto get_list = [&](int species, int kk, int si, int so, int depth) {
		const int gi = group_of.at({species, kk, si, so});
		auto key = std::make_pair(gi, depth);
		auto it = list_id.find(key);
		if (it != list_id.end())
			return it->second;
		const int id = (int)in->lists.size();
		const Species &s = in->sp[species];
		ListData ld;
		ld.species = species;
		ld.k = kk;
		ld.E_bra = s.subs[so].E;
		ld.m2_bra = s.subs[so].m2;
		ld.E_ket = s.subs[si].E;
		ld.m2_ket = s.subs[si].m2;
		ld.depth = depth;
		ld.entries = std::move(group_lists[gi].at(depth));
		in->lists.push_back(std::move(ld));
		list_id[key] = id;
		return id;
	};
	auto get_v = [&](const VKey &key) {
		auto it = v_id.find(key);
		if (it != v_id.end())
			return it->second;
		const int id = (int)in->vkeys.size();
		in->vkeys.push_back(key);
		v_id[key] = id;
		return id;
	};
	for (const Req &r : reqs) {
		const int pin = in->block_sub_p[r.bi],
			  pout = in->block_sub_p[r.bo];
		const int nin = in->block_sub_n[r.bi],
			  nout = in->block_sub_n[r.bo];
		const Species &P = in->sp[0], &N = in->sp[1];
		const int dEp = P.subs[pout].E - P.subs[pin].E;
		const int dMp = P.subs[pout].m2 - P.subs[pin].m2;
		const int dEn = N.subs[nout].E - N.subs[nin].E;
		const int dMn = N.subs[nout].m2 - N.subs[nin].m2;
		const int64_t np_in = in->blocks[6 * r.bi + 4];
		const int64_t nn_in = in->blocks[6 * r.bi + 5];
		if (r.type == 1) {
			auto &gl = group_lists[group_of.at({1, r.kn, nin, nout})];
			std::vector<int> depths;
			for (auto &kv : gl)
				depths.push_back(kv.first);
			for (int d : depths) {
				const int L = get_list(1, r.kn, nin, nout, d);
				const int V = get_v({0, r.kn, 0, 0, 0, d, dEn, dMn});
				ncsm_instruction_t ins{1, r.rank, r.bi, r.bo, L, -1, V, 0};
				in->instr.push_back(ins);
				in->instr_mults.push_back(
					(int64_t)(in->lists[L].entries.size() / 3) * np_in);
			}
		} else if (r.type == 2) {
			auto &gl = group_lists[group_of.at({0, r.kp, pin, pout})];
			std::vector<int> depths;
			for (auto &kv : gl)
				depths.push_back(kv.first);
			for (int d : depths) {
				const int L = get_list(0, r.kp, pin, pout, d);
				const int V = get_v({r.kp, 0, d, dEp, dMp, 0, 0, 0});
				ncsm_instruction_t ins{2, r.rank, r.bi, r.bo, -1, L, V, 0};
				in->instr.push_back(ins);
				in->instr_mults.push_back(
					(int64_t)(in->lists[L].entries.size() / 3) * nn_in);
			}
		} else {
			const int gn = group_of.at({1, r.kn, nin, nout});
			const int gp = group_of.at({0, r.kp, pin, pout});
			std::vector<int> dns, dps;
			for (auto &kv : group_lists[gn])
				dns.push_back(kv.first);
			for (auto &kv : group_lists[gp])
				dps.push_back(kv.first);
			for (int dn : dns)
				for (int dp : dps) {
					const int Ln = get_list(1, r.kn, nin, nout, dn);
					const int Lp = get_list(0, r.kp, pin, pout, dp);
					const int V = get_v({r.kp, r.kn, dp, dEp, dMp,
							     dn, dEn, dMn});
					ncsm_instruction_t ins{3, r.rank, r.bi, r.bo,
							       Ln, Lp, V, 0};
					in->instr.push_back(ins);
					in->instr_mults.push_back(
						(int64_t)(in->lists[Ln].entries.size() / 3) *
						(int64_t)(in->lists[Lp].entries.size() / 3));
				}
		}
	}
	// flatten lists
	in->list_offsets.push_back(0);
	for (const ListData &ld : in->lists) {
		in->list_entries.insert(in->list_entries.end(),
					ld.entries.begin(), ld.entries.end());
		in->list_offsets.push_back((int64_t)in->list_entries.size() / 3);
		in->list_settings.insert(in->list_settings.end(),
					 {ld.species, ld.k, ld.E_bra, ld.m2_bra,
					  ld.E_ket, ld.m2_ket, ld.depth});
	}
	for (ListData &ld : in->lists)
		std::vector<int32_t>().swap(ld.entries);

	clk.mark("instructions, flattened lists");
	// pass 4: matrix-element blocks
	const size_t nv = in->vkeys.size();
	std::vector<ConnList> cp(nv), cn(nv);
	in->v_offsets.push_back(0);
	for (size_t v = 0; v < nv; v++) {
		const VKey &k = in->vkeys[v];
		cp[v] = conn_list(*in, 0, k.kp, k.dp, k.dEp, k.dMp);
		cn[v] = conn_list(*in, 1, k.kn, k.dn, k.dEn, k.dMn);
		const int64_t Dp = k.kp ? (int64_t)cp[v].pairs.size() : 0;
		const int64_t Dn = k.kn ? (int64_t)cn[v].pairs.size() : 0;
		const int64_t n = (Dn == 0) ? Dp : (Dp == 0 ? Dn : Dn * Dp);
		in->v_offsets.push_back(in->v_offsets.back() + n);
		in->v_shapes.insert(in->v_shapes.end(), {Dn, Dp, k.kp + k.kn});
	}
	clk.mark("connection pairs of the element blocks");
	in->v_elements.resize(in->v_offsets.back());
	{
		// transparent huge pages for the element array: its first
		// touch (the writes below) then faults in 2 MiB pages, not 4 KiB
		const uintptr_t a = ((uintptr_t)in->v_elements.data() + 4095) & ~(uintptr_t)4095;
		const uintptr_t b = (uintptr_t)(in->v_elements.data() + in->v_elements.size());
		if (b > a + (2u << 20))
			(void)madvise((void *)a, b - a, MADV_HUGEPAGE);
	}
	clk.mark("element array");
	if (gpu) {
		// one device thread per element: the pairs of every block's
		// proton and neutron parts, packed
		std::vector<ncsm::ElemBlock> eb(nv);
		int64_t npairs = 0;
		for (size_t v = 0; v < nv; v++) {
			const VKey &k = in->vkeys[v];
			ncsm::ElemBlock &b = eb[v];
			b.offset = in->v_offsets[v];
			b.pair_first = npairs;
			b.np = (int32_t)cp[v].pairs.size();
			b.nn = (int32_t)cn[v].pairs.size();
			b.kp = k.kp;
			b.kn = k.kn;
			b.sigma = k.kp + k.kn == 2 ? c.sigma2 : c.sigma3;
			npairs += b.np + b.nn;
		}
		std::vector<ncsm::Packed> pairs(2 * npairs);
#pragma omp parallel for schedule(dynamic, 16)
		for (long v = 0; v < (long)nv; v++) {
			const VKey &k = in->vkeys[v];
			ncsm::Packed *dst = pairs.data() + 2 * eb[v].pair_first;
			for (int part = 0; part < 2; part++) {
				const ConnList &L = part == 0 ? cp[v] : cn[v];
				const int kk = part == 0 ? k.kp : k.kn;
				for (const auto &pr : L.pairs) {
					*dst++ = kk ? pack(L.ann->states[pr.first]) : 0;
					*dst++ = kk ? pack(L.cre->states[pr.second]) : 0;
				}
			}
		}
		std::vector<int32_t> orbital_e;
		for (const Orbital &o : in->orbs)
			orbital_e.push_back(o.e);
		clk.mark("element pairs (host)");
		std::string err;
		const int rc = ncsm::gpu_elements(eb, pairs, orbital_e, c.seed, c.hw,
						  in->v_elements.data(),
						  (int64_t)in->v_elements.size(), err);
		clk.mark("elements (device)");
		if (rc) {
			fprintf(stderr, "ncsm_inputs_build_gpu: matrix elements: %s\n", err.c_str());
			delete in;
			return nullptr;
		}
		return in;
	}
	for (size_t v = 0; v < nv; v++) {
		const VKey &k = in->vkeys[v];
		const int rank = k.kp + k.kn;
		const double sigma = rank == 2 ? c.sigma2 : c.sigma3;
		const ConnList &P = cp[v], &Nl = cn[v];
		const int64_t Dn = (int64_t)Nl.pairs.size();
		const int64_t npair = (int64_t)P.pairs.size();
		double *dst = in->v_elements.data() + in->v_offsets[v];
		const ncsm_inputs *cin = in;
#pragma omp parallel for collapse(2) schedule(static)
		for (int64_t mp = 0; mp < npair; mp++) {
			for (int64_t mn = 0; mn < Dn; mn++) {
				const Tuple *pa = k.kp ? &P.ann->states[P.pairs[mp].first] : nullptr;
				const Tuple *pc = k.kp ? &P.cre->states[P.pairs[mp].second] : nullptr;
				const Tuple *na = k.kn ? &Nl.ann->states[Nl.pairs[mn].first] : nullptr;
				const Tuple *nc = k.kn ? &Nl.cre->states[Nl.pairs[mn].second] : nullptr;
				int32_t bra[6], ket[6];
				const int cnt = fill_global(bra, pc, k.kp, nc, k.kn);
				fill_global(ket, pa, k.kp, na, k.kn);
				double val;
				if (rank == 1) {
					val = (bra[0] == ket[0])
						? cin->cfg.hw * (cin->orbs[bra[0] >> 1].e + 1.5)
						: 0.0;
				} else
					val = sigma * keyed_gaussian(cin->cfg.seed, rank,
								     bra, ket, cnt);
				dst[mn + Dn * mp] = val;
			}
		}
	}
	clk.mark("elements (host)");
	return in;
}

}  // namespace

void ncsm_inputs_free(ncsm_inputs_t *inputs) { delete inputs; }

int64_t ncsm_dimension(const ncsm_inputs_t *in) { return in->dimension; }
int64_t ncsm_num_blocks(const ncsm_inputs_t *in)
{
	return (int64_t)in->blocks.size() / 6;
}
const int64_t *ncsm_blocks(const ncsm_inputs_t *in) { return in->blocks.data(); }
int64_t ncsm_num_lists(const ncsm_inputs_t *in)
{
	return (int64_t)in->list_offsets.size() - 1;
}
const int64_t *ncsm_list_offsets(const ncsm_inputs_t *in)
{
	return in->list_offsets.data();
}
const int32_t *ncsm_list_entries(const ncsm_inputs_t *in)
{
	return in->list_entries.data();
}
const int32_t *ncsm_list_settings(const ncsm_inputs_t *in)
{
	return in->list_settings.data();
}
int64_t ncsm_num_matrix_blocks(const ncsm_inputs_t *in)
{
	return (int64_t)in->vkeys.size();
}
const int64_t *ncsm_matrix_offsets(const ncsm_inputs_t *in)
{
	return in->v_offsets.data();
}
const double *ncsm_matrix_elements(const ncsm_inputs_t *in)
{
	return in->v_elements.data();
}
const int64_t *ncsm_matrix_shapes(const ncsm_inputs_t *in)
{
	return in->v_shapes.data();
}
int64_t ncsm_num_instructions(const ncsm_inputs_t *in)
{
	return (int64_t)in->instr.size();
}
const ncsm_instruction_t *ncsm_instructions(const ncsm_inputs_t *in)
{
	return in->instr.data();
}
int32_t ncsm_num_orbitals(const ncsm_inputs_t *in)
{
	return (int32_t)in->orbs.size();
}
const int32_t *ncsm_orbitals(const ncsm_inputs_t *in)
{
	return in->orbital_table.data();
}

int64_t ncsm_subbasis_states(const ncsm_inputs_t *in, int64_t block,
			     int species, int32_t *out)
{
	if (block < 0 || block >= (int64_t)in->block_sub_p.size() ||
	    species < 0 || species > 1)
		return -1;
	const Species &s = in->sp[species];
	const SubBasis &sb = s.subs[species == 0 ? in->block_sub_p[block]
						 : in->block_sub_n[block]];
	if (out)
		for (size_t i = 0; i < sb.dets.size(); i++)
			for (int j = 0; j < s.k; j++)
				out[i * s.k + j] = sb.dets[i][j];
	return (int64_t)sb.dets.size();
}

double ncsm_element_value(const ncsm_inputs_t *in, int rank,
			  const int32_t *bra, const int32_t *ket)
{
	if (rank < 1 || rank > 3)
		return 0.0;
	int mb = 0, mk = 0, eb = 0, ek = 0, pb = 0, pk = 0;
	const int norb = (int)in->orbs.size();
	for (int i = 0; i < rank; i++) {
		if (bra[i] < 0 || ket[i] < 0 || (bra[i] >> 1) >= norb ||
		    (ket[i] >> 1) >= norb)
			return 0.0;
		mb += in->orbs[bra[i] >> 1].m2;
		mk += in->orbs[ket[i] >> 1].m2;
		eb += in->orbs[bra[i] >> 1].e;
		ek += in->orbs[ket[i] >> 1].e;
		pb += 1 - (bra[i] & 1);
		pk += 1 - (ket[i] & 1);
	}
	if (mb != mk || ((eb - ek) & 1) || pb != pk)
		return 0.0;
	if (rank == 1)
		return bra[0] == ket[0]
			       ? in->cfg.hw * (in->orbs[bra[0] >> 1].e + 1.5)
			       : 0.0;
	const double sigma = rank == 2 ? in->cfg.sigma2 : in->cfg.sigma3;
	return sigma * keyed_gaussian(in->cfg.seed, rank, bra, ket, rank);
}

int ncsm_inputs_write(const ncsm_inputs_t *in, const char *directory)
{
	const std::string dir(directory);
	if (!make_dir(dir) || !make_dir(dir + "/index_lists") ||
	    !make_dir(dir + "/matrix_elements"))
		return -1;
	const int64_t nb = ncsm_num_blocks(in);
	const int64_t nl = ncsm_num_lists(in);
	const int64_t nv = ncsm_num_matrix_blocks(in);
	const int64_t first_list = nb + 1, first_v = nb + nl + 1;
	FILE *f = fopen((dir + "/comb.txt").c_str(), "w");
	if (!f)
		return -1;
	fprintf(f, "*** Combination table: synthetic Z=%d N=%d Nmax=%d "
		   "2M=%d seed=%llu ***\n",
		in->cfg.num_protons, in->cfg.num_neutrons, in->cfg.nmax,
		in->cfg.twice_m, (unsigned long long)in->cfg.seed);
	fprintf(f, "*** mp-states ***\n");
	fprintf(f, "#  Ep  Mp  En  Mn      Np      Nn     dim\n");
	for (int64_t b = 0; b < nb; b++) {
		const int64_t *B = &in->blocks[6 * b];
		fprintf(f, "%4lld %4lld %4lld %4lld %7lld %7lld %9lld  # ARRAYMP: %lld= %lld\n",
			(long long)B[0], (long long)B[1], (long long)B[2],
			(long long)B[3], (long long)B[4], (long long)B[5],
			(long long)(B[4] * B[5]), (long long)(b + 1),
			(long long)(B[4] * B[5] * 8));
	}
	fprintf(f, "*** Conn lists ***\n");
	for (int64_t l = 0; l < nl; l++) {
		const int32_t *S = &in->list_settings[7 * l];
		const int64_t len = in->list_offsets[l + 1] - in->list_offsets[l];
		fprintf(f, "%-4s %4d %4d %4d %4d %4d %9lld  # ARRAY: %lld= %lld\n",
			type_string(S[1], S[0] == 0 ? 'p' : 'n'), S[2], S[3], S[4],
			S[5], S[6], (long long)len, (long long)(first_list + l),
			(long long)(len * 12));
	}
	fprintf(f, "*** Matrix-elements V (cross p-n) ***\n");
	for (int64_t v = 0; v < nv; v++) {
		const VKey &k = in->vkeys[v];
		if (!(k.kp && k.kn))
			continue;
		const int64_t Dn = in->v_shapes[3 * v], Dp = in->v_shapes[3 * v + 1];
		fprintf(f, "%-4s %4d %4d %4d %-4s %4d %4d %4d %9lld %9lld %12lld  # ARRAY: %lld= %lld\n",
			type_string(k.kp, 'p'), k.dEp, k.dMp, k.dp,
			type_string(k.kn, 'n'), k.dEn, k.dMn, k.dn, (long long)Dp,
			(long long)Dn, (long long)(Dn * Dp),
			(long long)(first_v + v), (long long)(Dn * Dp * 8));
	}
	fprintf(f, "*** Matrix-elements V (same p/n) ***\n");
	for (int64_t v = 0; v < nv; v++) {
		const VKey &k = in->vkeys[v];
		if (k.kp && k.kn)
			continue;
		const int64_t D = k.kp ? in->v_shapes[3 * v + 1] : in->v_shapes[3 * v];
		if (k.kp)
			fprintf(f, "%-4s %4d %4d %4d %9lld %12lld  # ARRAY: %lld= %lld\n",
				type_string(k.kp, 'p'), k.dEp, k.dMp, k.dp,
				(long long)D, (long long)D,
				(long long)(first_v + v), (long long)(D * 8));
		else
			fprintf(f, "%-4s %4d %4d %4d %9lld %12lld  # ARRAY: %lld= %lld\n",
				type_string(k.kn, 'n'), k.dEn, k.dMn, k.dn,
				(long long)D, (long long)D,
				(long long)(first_v + v), (long long)(D * 8));
	}
	fprintf(f, "*** Calculation blocks ***\n");
	for (size_t i = 0; i < in->instr.size(); i++) {
		const ncsm_instruction_t &I = in->instr[i];
		if (I.type == 3)
			fprintf(f, "%12lld  # CALCBLOCK:%d: %d, %d, %lld, %lld, %lld\n",
				(long long)in->instr_mults[i], I.num_particles,
				I.block_in + 1, I.block_out + 1,
				(long long)(first_list + I.neutron_list),
				(long long)(first_list + I.proton_list),
				(long long)(first_v + I.matrix_block));
		else
			fprintf(f, "%12lld  # CALCBLOCK:%d: %d, %d, %lld, %lld\n",
				(long long)in->instr_mults[i], I.num_particles,
				I.block_in + 1, I.block_out + 1,
				(long long)(first_list +
					    (I.type == 1 ? I.neutron_list : I.proton_list)),
				(long long)(first_v + I.matrix_block));
	}
	fclose(f);

	f = fopen((dir + "/evaluation_order").c_str(), "w");
	if (!f)
		return -1;
	for (const ncsm_instruction_t &I : in->instr) {
		if (I.type == 3)
			fprintf(f, "BLOCK: %d %d %lld %lld %lld\n", I.block_in + 1,
				I.block_out + 1,
				(long long)(first_list + I.neutron_list),
				(long long)(first_list + I.proton_list),
				(long long)(first_v + I.matrix_block));
		else
			fprintf(f, "BLOCK: %d %d %lld %lld\n", I.block_in + 1,
				I.block_out + 1,
				(long long)(first_list +
					    (I.type == 1 ? I.neutron_list : I.proton_list)),
				(long long)(first_v + I.matrix_block));
	}
	fclose(f);

	for (int64_t l = 0; l < nl; l++) {
		char name[64];
		snprintf(name, sizeof name, "/index_lists/index_list_%lld",
			 (long long)(first_list + l));
		FILE *g = fopen((dir + name).c_str(), "wb");
		if (!g)
			return -1;
		const int64_t a = in->list_offsets[l], b = in->list_offsets[l + 1];
		if (b > a && fwrite(&in->list_entries[3 * a], 12, b - a, g) !=
				     (size_t)(b - a)) {
			fclose(g);
			return -1;
		}
		fclose(g);
	}
	for (int64_t v = 0; v < nv; v++) {
		char name[64];
		snprintf(name, sizeof name, "/matrix_elements/%lld_matrix_elements",
			 (long long)(first_v + v));
		FILE *g = fopen((dir + name).c_str(), "wb");
		if (!g)
			return -1;
		const uint64_t dn = (uint64_t)in->v_shapes[3 * v];
		const uint64_t dp = (uint64_t)in->v_shapes[3 * v + 1];
		const int64_t a = in->v_offsets[v], b = in->v_offsets[v + 1];
		bool ok = fwrite(&dn, 8, 1, g) == 1 && fwrite(&dp, 8, 1, g) == 1;
		if (ok && b > a)
			ok = fwrite(&in->v_elements[a], 8, b - a, g) == (size_t)(b - a);
		fclose(g);
		if (!ok)
			return -1;
	}
	return 0;
}

}  // extern "C"
