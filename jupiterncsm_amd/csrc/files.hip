// files.hip - the file-based boundary: comb.txt, the BLOCK: evaluation
// order, index_list_<id>, <id>_matrix_elements and vec_<id> directories,
// read and written with the reference's layouts.
#include "bacchus_internal.h"

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <array>
#include <map>
#include <string>
#include <vector>

namespace bacchus {
namespace {

std::vector<std::string> split_words(const std::string &row)
{
	// whitespace = ' ', '\n', '\t' (src/Utilities/string_tools/
	// string_tools.c:222-226)
	std::vector<std::string> words;
	size_t i = 0;
	while (i < row.size()) {
		while (i < row.size() && (row[i] == ' ' || row[i] == '\n' || row[i] == '\t'))
			i++;
		size_t j = i;
		while (j < row.size() && !(row[j] == ' ' || row[j] == '\n' || row[j] == '\t'))
			j++;
		if (j > i)
			words.push_back(row.substr(i, j - i));
		i = j;
	}
	return words;
}

// digits up to '=' (combination_table.c:715-723)
int64_t id_of(const std::string &w)
{
	int64_t id = 0;
	for (char c : w) {
		if (c == '=')
			break;
		id = 10 * id + (c - '0');
	}
	return id;
}

bool read_lines(const char *path, std::vector<std::string> &lines)
{
	FILE *f = fopen(path, "r");
	if (!f)
		return false;
	char *row = nullptr;
	size_t cap = 0;
	ssize_t n;
	while ((n = getline(&row, &cap, f)) >= 0)
		lines.emplace_back(row, (size_t)n);
	free(row);
	fclose(f);
	return true;
}

// One CALCBLOCK row (calculation_blocks.c:31-84); l2 = -1 when absent.
struct CalcBlock {
	int64_t num_particles, in, out, l1, l2, v, mults;
};

struct CombTable {
	std::vector<int64_t> block_ids, block_np, block_nn;
	std::map<int64_t, int64_t> block_index;  // id -> dense
	std::map<int64_t, char> list_type;       // id -> 'n' / 'p'
	std::map<int64_t, int> matrix_rank;      // id -> 1, 2, 3 body
	// bytes of every array by id, as get_array_sizes computes them
	// (combination_table.c:384-401): blocks 8 Np Nn, lists 12 length,
	// element blocks 8 x their length (matrix_block_setting.c:3-25)
	std::map<int64_t, int64_t> array_size;
	std::vector<CalcBlock> calc;
};

// Sections by title (combination_table.c:419-640): rows of a section run
// from the first row containing the title to the next row containing "***".
int parse_comb(const char *path, CombTable &t)
{
	std::vector<std::string> lines;
	if (!read_lines(path, lines))
		return BACCHUS_FAIL("Could not open \"%s\". %s", path, strerror(errno));
	auto section = [&](const char *title, auto &&fn) {
		size_t i = 0;
		while (i < lines.size() && lines[i].find(title) == std::string::npos)
			i++;
		for (i++; i < lines.size(); i++) {
			if (lines[i].find("***") != std::string::npos)
				break;
			fn(lines[i]);
		}
	};
	int bad = 0;
	section("mp-states", [&](const std::string &row) {
		if (row.find("ARRAYMP:") == std::string::npos)
			return;
		auto w = split_words(row);
		if (w.size() < 10) {
			bad = 1;
			return;
		}
		const int64_t id = id_of(w[9]);
		t.block_index[id] = (int64_t)t.block_ids.size();
		t.block_ids.push_back(id);
		t.block_np.push_back(atoll(w[4].c_str()));
		t.block_nn.push_back(atoll(w[5].c_str()));
		t.array_size[id] = 8 * t.block_np.back() * t.block_nn.back();
	});
	section("Conn lists", [&](const std::string &row) {
		if (row.find("ARRAY:") == std::string::npos)
			return;
		auto w = split_words(row);
		if (w.size() < 10 || (w[0][0] != 'n' && w[0][0] != 'p')) {
			bad = 1;
			return;
		}
		t.list_type[id_of(w[9])] = w[0][0];
		t.array_size[id_of(w[9])] = 12 * atoll(w[6].c_str());
	});
	// cross p-n rows: type = word 4 + word 0, id = word 13; same p/n rows:
	// type = word 0, id = word 8 (combination_table.c:535-637)
	section("Matrix-elements V (cross p-n)", [&](const std::string &row) {
		if (row.find("ARRAY:") == std::string::npos)
			return;
		auto w = split_words(row);
		if (w.size() < 14) {
			bad = 1;
			return;
		}
		t.matrix_rank[id_of(w[13])] = (int)(w[0].size() + w[4].size());
		t.array_size[id_of(w[13])] = 8 * atoll(w[8].c_str()) * atoll(w[9].c_str());
	});
	section("Matrix-elements V (same p/n)", [&](const std::string &row) {
		if (row.find("ARRAY:") == std::string::npos)
			return;
		auto w = split_words(row);
		if (w.size() < 9) {
			bad = 1;
			return;
		}
		t.matrix_rank[id_of(w[8])] = (int)w[0].size();
		t.array_size[id_of(w[8])] = 8 * atoll(w[4].c_str());
	});
	// CALCBLOCK rows anywhere in the file; the multiplication count is the
	// word two words before "CALCBLOCK:" (calculation_blocks.c:47-76)
	for (const std::string &row : lines) {
		const size_t at = row.find("CALCBLOCK:");
		if (at == std::string::npos)
			continue;
		CalcBlock c{};
		long long a[6];
		const char *cb = row.c_str() + at;
		const int n6 = sscanf(cb, "CALCBLOCK:%lld: %lld, %lld, %lld, %lld, %lld", &a[0], &a[1],
				      &a[2], &a[3], &a[4], &a[5]);
		if (n6 == 6) {
			c = {a[0], a[1], a[2], a[3], a[4], a[5], 0};
		} else if (sscanf(cb, "CALCBLOCK:%lld: %lld, %lld, %lld, %lld", &a[0], &a[1], &a[2],
				  &a[3], &a[4]) == 5) {
			c = {a[0], a[1], a[2], a[3], -1, a[4], 0};
		} else {
			bad = 2;
			continue;
		}
		const auto before = split_words(row.substr(0, at));
		if (before.size() < 2) {
			bad = 2;
			continue;
		}
		c.mults = atoll(before[before.size() - 2].c_str());
		t.calc.push_back(c);
	}
	if (bad == 2)
		return BACCHUS_FAIL("CALCBLOCK row in %s is not formatted correctly", path);
	if (bad)
		return BACCHUS_FAIL("malformed ARRAYMP/ARRAY row in %s", path);
	if (t.block_ids.empty())
		return BACCHUS_FAIL("no basis blocks (ARRAYMP rows) in %s", path);
	return 0;
}

int read_file(const std::string &path, std::vector<char> &buf)
{
	FILE *f = fopen(path.c_str(), "rb");
	if (!f)
		return BACCHUS_FAIL("Could not open file %s. %s", path.c_str(), strerror(errno));
	fseek(f, 0, SEEK_END);
	const long n = ftell(f);
	fseek(f, 0, SEEK_SET);
	buf.resize(n > 0 ? (size_t)n : 0);
	if (n > 0 && fread(buf.data(), 1, (size_t)n, f) != (size_t)n) {
		fclose(f);
		return BACCHUS_FAIL("Could not read %s", path.c_str());
	}
	fclose(f);
	return 0;
}

// The calculation-block order of Mars (src/Mars/programs/mars.c:92-170):
// blocks grouped by particle number, the 1-body ones dropped (and the
// 3-body ones in two-nucleon mode), then seven stable sorts whose last key
// is the primary one: heaviest first, then the size of the largest array,
// the ids of the five arrays ordered by size.
std::vector<CalcBlock> mars_order(const CombTable &t, bool two_nucleon_only)
{
	constexpr uint64_t kNoIndex = ~0ull;  // no_index, global_constants.h:8
	std::vector<CalcBlock> v;
	for (int np = 1; np <= 3; np++)
		for (const CalcBlock &c : t.calc)
			if (c.num_particles == np)
				v.push_back(c);
	std::vector<CalcBlock> need;
	for (const CalcBlock &c : v)
		if (c.num_particles == 2 || (!two_nucleon_only && c.num_particles == 3))
			need.push_back(c);
	auto size_of = [&](uint64_t id) -> uint64_t {
		if (id == kNoIndex)
			return 0;
		auto it = t.array_size.find((int64_t)id);
		return it == t.array_size.end() ? 0 : (uint64_t)it->second;
	};
	// the five arrays ordered by size with mars.c's own compare-and-swap
	// sequence (quint_sort, :322-340)
	auto arrays = [&](const CalcBlock &c) {
		std::array<uint64_t, 5> a = {(uint64_t)c.in, (uint64_t)c.out, (uint64_t)c.l1,
					     c.l2 < 0 ? kNoIndex : (uint64_t)c.l2, (uint64_t)c.v};
		for (int i = 1; i < 5; i++)
			for (int j = 0; j < i; j++)
				if (size_of(a[j]) > size_of(a[j + 1]))
					std::swap(a[j], a[j + 1]);
		return a;
	};
	auto by = [&](auto key) {
		std::stable_sort(need.begin(), need.end(), [&](const CalcBlock &x, const CalcBlock &y) {
			return key(x) < key(y);
		});
	};
	by([&](const CalcBlock &c) { return arrays(c)[0]; });
	by([&](const CalcBlock &c) { return arrays(c)[1]; });
	by([&](const CalcBlock &c) { return arrays(c)[2]; });
	by([&](const CalcBlock &c) { return arrays(c)[3]; });
	by([&](const CalcBlock &c) { return arrays(c)[4]; });
	by([&](const CalcBlock &c) { return size_of(arrays(c)[4]); });
	by([&](const CalcBlock &c) { return (uint64_t)(-c.mults); });
	return need;
}

// One instruction with the reference's file ids: basis blocks as dense
// indices, index lists and the element block by id (0 = no list).
struct IdInstruction {
	int32_t type;
	int32_t block_in, block_out;
	int64_t neutron_list, proton_list, matrix_block;
};

// Reads the index lists (index_list.c:69-97) and element blocks
// (matrix_block.c:22-62) the instructions name, checks the memory budget the
// way the reference's memory manager does (a single instruction needing more
// than maximum_loaded_memory is an error, memory_manager.c:354-365) and
// builds the device operator.  maximum_loaded_memory = 0 means no budget.
int create_from_ids(const std::vector<int64_t> &block_ids, const std::vector<int64_t> &block_np,
		    const std::vector<int64_t> &block_nn, const std::vector<IdInstruction> &ins,
		    const std::map<int64_t, int> &matrix_rank, const char *index_dir,
		    const char *matrix_dir, size_t maximum_loaded_memory, int precision,
		    bacchus_hip_operator_t **op)
{
	std::vector<bacchus_hip_instruction_t> instr;
	std::map<int64_t, int32_t> list_index, v_index;
	std::vector<int64_t> list_ids, v_ids;
	auto dense = [](std::map<int64_t, int32_t> &m, std::vector<int64_t> &ids, int64_t id) {
		auto it = m.find(id);
		if (it != m.end())
			return it->second;
		const int32_t k = (int32_t)ids.size();
		ids.push_back(id);
		m[id] = k;
		return k;
	};
	const int64_t nb = (int64_t)block_ids.size();
	for (const IdInstruction &w : ins) {
		if (w.block_in < 0 || w.block_in >= nb || w.block_out < 0 || w.block_out >= nb)
			return BACCHUS_FAIL("instruction names an unknown basis block");
		if (w.type < 1 || w.type > 3)
			return BACCHUS_FAIL("instruction has unknown type %d", w.type);
		bacchus_hip_instruction_t I{};
		I.type = w.type;
		I.block_in = w.block_in;
		I.block_out = w.block_out;
		I.neutron_list = w.type != 2 ? dense(list_index, list_ids, w.neutron_list) : -1;
		I.proton_list = w.type != 1 ? dense(list_index, list_ids, w.proton_list) : -1;
		I.matrix_block = dense(v_index, v_ids, w.matrix_block);
		instr.push_back(I);
	}
	// index lists
	std::vector<int64_t> list_offsets{0};
	std::vector<int32_t> list_entries;
	std::vector<char> buf;
	for (int64_t id : list_ids) {
		char name[64];
		snprintf(name, sizeof name, "/index_list_%lld", (long long)id);
		if (read_file(std::string(index_dir) + name, buf))
			return 1;
		if (buf.size() % 12 != 0)
			return BACCHUS_FAIL("index list %lld is not a whole number of "
					    "triples", (long long)id);
		const size_t n = buf.size() / 4;
		const size_t old = list_entries.size();
		list_entries.resize(old + n);
		if (n)
			memcpy(&list_entries[old], buf.data(), buf.size());
		list_offsets.push_back((int64_t)list_entries.size() / 3);
	}
	// matrix elements: size_t Dn, size_t Dp, double[n] with n = Dn*Dp, or Dp
	// if Dn = 0, or Dn if Dp = 0
	std::vector<int64_t> v_offsets{0}, v_dn;
	std::vector<int32_t> v_force;
	std::vector<double> elements;
	for (int64_t id : v_ids) {
		char name[64];
		snprintf(name, sizeof name, "/%lld_matrix_elements", (long long)id);
		if (read_file(std::string(matrix_dir) + name, buf))
			return 1;
		if (buf.size() < 16)
			return BACCHUS_FAIL("Could not read neutron_matrix_dimension from "
					    "matrix file %lld", (long long)id);
		uint64_t dn, dp;
		memcpy(&dn, buf.data(), 8);
		memcpy(&dp, buf.data() + 8, 8);
		const uint64_t n = dn == 0 ? dp : (dp == 0 ? dn : dn * dp);
		if (buf.size() < 16 + 8 * n)
			return BACCHUS_FAIL("Could not read matrix elements form matrix "
					    "file %lld", (long long)id);
		const size_t old = elements.size();
		elements.resize(old + n);
		if (n)
			memcpy(&elements[old], buf.data() + 16, 8 * n);
		v_offsets.push_back((int64_t)elements.size());
		v_dn.push_back((int64_t)dn);
		auto r = matrix_rank.find(id);
		v_force.push_back(r == matrix_rank.end() ? 2 : std::min(3, std::max(1, r->second)));
	}
	// the budget, per instruction: its vector blocks, element block and
	// index lists (the arrays get_size_of_unloaded_arrays adds up,
	// memory_manager.c:328-352)
	if (maximum_loaded_memory) {
		for (size_t k = 0; k < instr.size(); k++) {
			const bacchus_hip_instruction_t &I = instr[k];
			size_t need = 8 * (size_t)(block_np[I.block_in] * block_nn[I.block_in]);
			if (I.block_out != I.block_in)
				need += 8 * (size_t)(block_np[I.block_out] * block_nn[I.block_out]);
			need += 8 * (size_t)(v_offsets[I.matrix_block + 1] - v_offsets[I.matrix_block]);
			for (int32_t l : {I.neutron_list, I.proton_list})
				if (l >= 0)
					need += 12 * (size_t)(list_offsets[l + 1] - list_offsets[l]);
			if (need > maximum_loaded_memory)
				return BACCHUS_FAIL("Block %zu needs %zu B to be loaded,But maximaly "
						    "allowed loaded memory is %zu.", k, need,
						    maximum_loaded_memory);
		}
	}
	bacchus_hip_problem_t p{};
	p.num_blocks = nb;
	p.block_np = block_np.data();
	p.block_nn = block_nn.data();
	p.num_lists = (int64_t)list_ids.size();
	p.list_offsets = list_offsets.data();
	p.list_entries = list_entries.data();
	p.num_matrix_blocks = (int64_t)v_ids.size();
	p.matrix_offsets = v_offsets.data();
	p.matrix_dn = v_dn.data();
	p.matrix_force = v_force.data();
	p.matrix_elements = elements.data();
	p.num_instructions = (int64_t)instr.size();
	p.instructions = instr.data();
	const int rc = bacchus_hip_operator_create(&p, precision, op);
	if (rc)
		return rc;
	(*op)->block_ids = block_ids;
	if (maximum_loaded_memory && (*op)->device_bytes > maximum_loaded_memory) {
		const size_t have = (*op)->device_bytes;
		bacchus_hip_operator_free(*op);
		*op = nullptr;
		return BACCHUS_FAIL("the operator needs %zu B of device memory, more than the "
				    "maximum_loaded_memory of %zu B (it is kept resident; there is "
				    "no out-of-core mode)", have, maximum_loaded_memory);
	}
	return 0;
}

}  // namespace
}  // namespace bacchus

using namespace bacchus;

extern "C" {

int bacchus_hip_operator_load(const char *combination_table_file,
			      int num_protons, int num_neutrons,
			      const char *evaluation_order_file,
			      const char *index_lists_base_directory,
			      const char *matrix_file_base_directory,
			      int precision, bacchus_hip_operator_t **op)
{
	(void)num_protons;
	(void)num_neutrons;
	if (!combination_table_file || !evaluation_order_file ||
	    !index_lists_base_directory || !matrix_file_base_directory || !op)
		return BACCHUS_FAIL("NULL argument");
	CombTable comb;
	if (parse_comb(combination_table_file, comb))
		return 1;
	// evaluation order (evaluation_order.c:175-243): rows with "BLOCK:",
	// "UNLOAD_" rows ignored, 6 words = NP, else N or P by list type
	std::vector<std::string> rows;
	if (!read_lines(evaluation_order_file, rows))
		return BACCHUS_FAIL("Could not open evaluation order file %s. %s",
				    evaluation_order_file, strerror(errno));
	std::vector<IdInstruction> instr;
	for (const std::string &row : rows) {
		const size_t at = row.find("BLOCK:");
		if (at == std::string::npos || row.find("UNLOAD_") != std::string::npos)
			continue;
		auto w = split_words(row.substr(at));
		if (w.size() != 5 && w.size() != 6)
			return BACCHUS_FAIL("evaluation order row \"%s\" is not formatted "
					    "correctly", row.c_str());
		IdInstruction I{};
		const int64_t bin = atoll(w[1].c_str()), bout = atoll(w[2].c_str());
		if (!comb.block_index.count(bin) || !comb.block_index.count(bout))
			return BACCHUS_FAIL("evaluation order row \"%s\" names an unknown "
					    "basis block", row.c_str());
		I.block_in = (int32_t)comb.block_index[bin];
		I.block_out = (int32_t)comb.block_index[bout];
		auto type_of = [&](int64_t id) -> char {
			auto it = comb.list_type.find(id);
			return it == comb.list_type.end() ? 0 : it->second;
		};
		if (w.size() == 6) {
			int64_t l1 = atoll(w[3].c_str()), l2 = atoll(w[4].c_str());
			if (!type_of(l1) || !type_of(l2))
				return BACCHUS_FAIL("row \"%s\" names an unknown index list",
						    row.c_str());
			if (type_of(l1) != 'n')
				std::swap(l1, l2);
			I.type = 3;
			I.neutron_list = l1;
			I.proton_list = l2;
			I.matrix_block = atoll(w[5].c_str());
		} else {
			const int64_t l = atoll(w[3].c_str());
			const char c = type_of(l);
			if (!c)
				return BACCHUS_FAIL("row \"%s\" names an unknown index list",
						    row.c_str());
			I.type = c == 'n' ? 1 : 2;
			I.neutron_list = c == 'n' ? l : 0;
			I.proton_list = c == 'n' ? 0 : l;
			I.matrix_block = atoll(w[4].c_str());
		}
		instr.push_back(I);
	}
	return create_from_ids(comb.block_ids, comb.block_np, comb.block_nn, instr,
			       comb.matrix_rank, index_lists_base_directory,
			       matrix_file_base_directory, 0, precision, op);
}

int bacchus_hip_operator_load_blocks(int64_t num_blocks, const int64_t *block_ids,
				     const int64_t *block_np, const int64_t *block_nn,
				     int64_t num_instructions,
				     const bacchus_hip_file_instruction_t *instructions,
				     const char *index_lists_base_directory,
				     const char *matrix_file_base_directory,
				     size_t maximum_loaded_memory, int precision,
				     bacchus_hip_operator_t **op)
{
	if (!op || !block_ids || !block_np || !block_nn || (num_instructions && !instructions) ||
	    !index_lists_base_directory || !matrix_file_base_directory || num_blocks <= 0 ||
	    num_instructions < 0)
		return BACCHUS_FAIL("NULL or invalid argument");
	*op = nullptr;
	std::vector<int64_t> ids(block_ids, block_ids + num_blocks);
	std::vector<int64_t> np(block_np, block_np + num_blocks), nn(block_nn, block_nn + num_blocks);
	std::map<int64_t, int64_t> index;
	for (int64_t b = 0; b < num_blocks; b++)
		index[ids[b]] = b;
	std::map<int64_t, int> rank;
	std::vector<IdInstruction> instr;
	for (int64_t k = 0; k < num_instructions; k++) {
		const bacchus_hip_file_instruction_t &f = instructions[k];
		if (!index.count(f.block_in) || !index.count(f.block_out))
			return BACCHUS_FAIL("instruction %lld names an unknown basis block",
					    (long long)k);
		IdInstruction I{};
		I.type = f.type;
		I.block_in = (int32_t)index[f.block_in];
		I.block_out = (int32_t)index[f.block_out];
		I.neutron_list = f.neutron_list;
		I.proton_list = f.proton_list;
		I.matrix_block = f.matrix_block;
		if (f.num_particles >= 1)
			rank[f.matrix_block] = f.num_particles;
		instr.push_back(I);
	}
	return create_from_ids(ids, np, nn, instr, rank, index_lists_base_directory,
			       matrix_file_base_directory, maximum_loaded_memory, precision, op);
}

int bacchus_hip_run_matrix_vector_multiplication(
	const char *output_vector_base_directory,
	const char *input_vector_base_directory, bacchus_hip_operator_t *op)
{
	if (!op || !output_vector_base_directory || !input_vector_base_directory)
		return BACCHUS_FAIL("NULL argument");
	const int64_t nb = (int64_t)op->block_np.size();
	const std::vector<int64_t> &ids = op->block_ids;
	const int64_t D = op->dimension;
	std::vector<double> x(D), y(D);
	auto io = [&](const char *dir, std::vector<double> &v, bool write) -> int {
		for (int64_t b = 0; b < nb; b++) {
			char name[64];
			snprintf(name, sizeof name, "/vec_%lld", (long long)ids[b]);
			const std::string path = std::string(dir) + name;
			FILE *f = fopen(path.c_str(), write ? "wb" : "rb");
			if (!f)
				return BACCHUS_FAIL("Could not open vector block file %s. %s",
						    path.c_str(), strerror(errno));
			const size_t n = (size_t)(op->block_off[b + 1] - op->block_off[b]);
			const size_t done = write ? fwrite(&v[op->block_off[b]], 8, n, f)
						  : fread(&v[op->block_off[b]], 8, n, f);
			fclose(f);
			if (done != n)
				return BACCHUS_FAIL("Could not %s elements of block %lld in %s",
						    write ? "write" : "read", (long long)ids[b], dir);
		}
		return 0;
	};
	if (io(input_vector_base_directory, x, false) ||
	    io(output_vector_base_directory, y, false))
		return 1;
	double *dx = nullptr, *dy = nullptr;
	BACCHUS_HIP_CHECK(hipMalloc(&dx, D * 8));
	if (hipMalloc(&dy, D * 8) != hipSuccess) {
		(void)hipFree(dx);
		return BACCHUS_FAIL("hipMalloc of %lld bytes failed", (long long)(D * 8));
	}
	int rc = 0;
	if (hipMemcpy(dx, x.data(), D * 8, hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemcpy(dy, y.data(), D * 8, hipMemcpyHostToDevice) != hipSuccess)
		rc = BACCHUS_FAIL("host to device copy failed");
	if (!rc)
		rc = bacchus_hip_operator_matvec(op, dx, dy, 1, nullptr);
	if (!rc && hipMemcpy(y.data(), dy, D * 8, hipMemcpyDeviceToHost) != hipSuccess)
		rc = BACCHUS_FAIL("device to host copy failed");
	(void)hipFree(dx);
	(void)hipFree(dy);
	if (rc)
		return rc;
	return io(output_vector_base_directory, y, true);
}

int bacchus_hip_write_evaluation_order(const char *combination_table_file,
				       int two_nucleon_forces_only, const char *output_path)
{
	if (!combination_table_file || !output_path)
		return BACCHUS_FAIL("NULL argument");
	CombTable t;
	if (int rc = parse_comb(combination_table_file, t))
		return rc;
	const auto order = mars_order(t, two_nucleon_forces_only != 0);
	FILE *f = fopen(output_path, "w");
	if (!f)
		return BACCHUS_FAIL("Could not open \"%s\" for writing. %s", output_path,
				    strerror(errno));
	for (const CalcBlock &c : order) {
		if (c.l2 < 0)
			fprintf(f, "BLOCK: %lld %lld %lld %lld\n", (long long)c.in, (long long)c.out,
				(long long)c.l1, (long long)c.v);
		else
			fprintf(f, "BLOCK: %lld %lld %lld %lld %lld\n", (long long)c.in,
				(long long)c.out, (long long)c.l1, (long long)c.l2, (long long)c.v);
	}
	if (fclose(f) != 0)
		return BACCHUS_FAIL("Could not write \"%s\"", output_path);
	return 0;
}

}  // extern "C"
