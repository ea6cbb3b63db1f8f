// A compiled C++ caller of libbcg.so through include/bcg.h only, as a C++
// driver of the reference would call it: no Python between the calls.
//
//   bcg_caller abi
//       prints the sizes and field offsets of the ABI's structs and the
//       default parameters (no GPU needed);
//   bcg_caller loop <dir>
//       runs the time loop split at its output point (bcg_sim_prepare, at
//       OutFlag bcg_mean_stress, bcg_sim_advance with the moved cells) from
//       <dir>/params.bin, env_params.bin (the structs' bytes), cells.bin (a
//       Restart file: t, dt, Cell[N]), wall.bin (BoxX * BoxY doubles) and
//       run.txt ("seed steps update_time output_time"); writes
//       out<k>_stress.bin (N x 5 doubles) and out<k>_moved.bin (Cell[N]) at
//       output step k, lineage.bin (int32 pairs), cells_final.bin (Cell[N])
//       and draws.bin (4 int32: the next rand() values).
//
// With -DBCG_WITH_REFERENCE_CELL (set by the build where the reference's
// Cell.h is on the include path) the record layout is also checked against
// the reference's struct at compile time.
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bcg.h"

#ifdef BCG_WITH_REFERENCE_CELL
#include "Cell.h"
static_assert(sizeof(Cell) == sizeof(bcg_cell), "bcg_cell is not the reference's Cell");
static_assert(offsetof(Cell, Length) == offsetof(bcg_cell, Length), "Length");
static_assert(offsetof(Cell, Velocity) == offsetof(bcg_cell, Velocity), "Velocity");
static_assert(offsetof(Cell, DynFric) == offsetof(bcg_cell, DynFric), "DynFric");
static_assert(offsetof(Cell, Type) == offsetof(bcg_cell, Type), "Type");
static_assert(offsetof(Cell, Ancestor) == offsetof(bcg_cell, Ancestor), "Ancestor");
#endif
static_assert(sizeof(bcg_cell) == 208, "the Restart record is 208 bytes");

namespace {

int fail(const char* what, int rc, bcg_ctx* ctx, bcg_env* env) {
    char buf[512] = "";
    if (ctx) bcg_last_error(ctx, buf, sizeof buf);
    if (env && !buf[0]) bcg_env_last_error(env, buf, sizeof buf);
    std::fprintf(stderr, "%s failed: rc=%d %s\n", what, rc, buf);
    return 1;
}

bool read_file(const std::string& path, std::vector<char>& out) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    out.resize(n > 0 ? (size_t)n : 0);
    bool ok = n <= 0 || std::fread(out.data(), 1, (size_t)n, f) == (size_t)n;
    std::fclose(f);
    return ok;
}

bool write_file(const std::string& path, const void* data, size_t n) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    bool ok = n == 0 || std::fwrite(data, 1, n, f) == n;
    return std::fclose(f) == 0 && ok;
}

#define OFF(T, f) std::printf("offset %s.%s %zu\n", #T, #f, offsetof(T, f))

int abi() {
    std::printf("sizeof bcg_cell %zu\nsizeof bcg_params %zu\nsizeof bcg_env_params %zu\n", sizeof(bcg_cell),
                sizeof(bcg_params), sizeof(bcg_env_params));
    std::printf("sizeof bcg_env_cell %zu\nsizeof bcg_aga %zu\nsizeof bcg_clock %zu\n", sizeof(bcg_env_cell),
                sizeof(bcg_aga), sizeof(bcg_clock));
    OFF(bcg_cell, time_p); OFF(bcg_cell, age_p); OFF(bcg_cell, Length); OFF(bcg_cell, Velocity);
    OFF(bcg_cell, StaFric); OFF(bcg_cell, Type); OFF(bcg_cell, Ancestor);
    OFF(bcg_params, BoxLength); OFF(bcg_params, BoxX); OFF(bcg_params, maxCells); OFF(bcg_params, planar);
    OFF(bcg_env_params, BoxLength); OFF(bcg_env_params, UpdateTime); OFF(bcg_env_params, maxO2_Agar);
    OFF(bcg_env_params, pAceAnant); OFF(bcg_clock, NextOutTime); OFF(bcg_clock, OutFlag);
    bcg_params p;
    bcg_env_params e;
    if (bcg_default_params(&p) != BCG_OK || bcg_env_default_params(&e) != BCG_OK) return 1;
    std::printf("default cellRadius %.17g L_divide %.17g BoxX %d BoxZ %d maxCells %d\n", p.cellRadius, p.L_divide,
                p.BoxX, p.BoxZ, p.maxCells);
    std::printf("default BoxZAgar %d maxLevelsMG %d UpdateTime %.17g maxCarbon %.17g\n", e.BoxZAgar, e.maxLevelsMG,
                e.UpdateTime, e.maxCarbon);
    return 0;
}

int loop(const std::string& dir) {
    std::vector<char> pb, eb, cb, wb, rb;
    if (!read_file(dir + "/params.bin", pb) || pb.size() != sizeof(bcg_params) ||
        !read_file(dir + "/env_params.bin", eb) || eb.size() != sizeof(bcg_env_params) ||
        !read_file(dir + "/cells.bin", cb) || cb.size() < 16 || (cb.size() - 16) % sizeof(bcg_cell) ||
        !read_file(dir + "/wall.bin", wb) || !read_file(dir + "/run.txt", rb)) {
        std::fprintf(stderr, "bad inputs in %s\n", dir.c_str());
        return 2;
    }
    bcg_params p;
    bcg_env_params ep;
    std::memcpy(&p, pb.data(), sizeof p);
    std::memcpy(&ep, eb.data(), sizeof ep);
    double t0, dt;
    std::memcpy(&t0, cb.data(), 8);
    std::memcpy(&dt, cb.data() + 8, 8);
    const int n0 = (int)((cb.size() - 16) / sizeof(bcg_cell));
    rb.push_back('\0');
    unsigned seed = 0;
    int steps = 0;
    double update_time = 0, output_time = 0;
    if (std::sscanf(rb.data(), "%u %d %lf %lf", &seed, &steps, &update_time, &output_time) != 4) return 2;
    if (wb.size() != sizeof(double) * (size_t)p.BoxX * p.BoxY) return 2;

    bcg_ctx* ctx = nullptr;
    bcg_env* env = nullptr;
    int rc = bcg_create(&p, &ctx);
    if (rc) return fail("bcg_create", rc, nullptr, nullptr);
    const size_t nh = (size_t)p.BoxX * p.refinementGridHeight * p.BoxY * p.refinementGridHeight;
    std::vector<double> height(nh, p.cellRadius), normal(3 * nh, 0.0);
    for (size_t i = 0; i < nh; i++) normal[3 * i + 2] = -1.0;
    std::vector<bcg_cell> cells(n0);
    std::memcpy(cells.data(), cb.data() + 16, cells.size() * sizeof(bcg_cell));
    int32_t n = 0;
    std::vector<int32_t> lineage, lin(2 * 4096);
    bcg_clock clk;
    int out_k = 0;
    if ((rc = bcg_set_cells(ctx, cells.data(), n0))) return fail("bcg_set_cells", rc, ctx, env);
    if ((rc = bcg_set_fields(ctx, height.data(), normal.data(), (const double*)wb.data(), nullptr, nullptr)))
        return fail("bcg_set_fields", rc, ctx, env);
    if ((rc = bcg_srand(ctx, seed))) return fail("bcg_srand", rc, ctx, env);
    if ((rc = bcg_env_create(&ep, ctx, &env))) return fail("bcg_env_create", rc, ctx, env);
    if ((rc = bcg_env_fields(env, BCG_ENV_INIT))) return fail("bcg_env_fields", rc, ctx, env);
    int32_t count = 0;
    if ((rc = bcg_env_nutrients(env, &count))) return fail("bcg_env_nutrients", rc, ctx, env);
    if ((rc = bcg_clock_init(&clk, t0, dt, output_time, update_time))) return fail("bcg_clock_init", rc, ctx, env);
    for (int k = 0; k < steps; k++) {
        if ((rc = bcg_sim_prepare(ctx, env, &clk))) return fail("bcg_sim_prepare", rc, ctx, env);
        const bool out = clk.OutFlag != 0;
        std::vector<bcg_cell> moved;
        if (out) {
            if ((rc = bcg_num_cells(ctx, &n))) return fail("bcg_num_cells", rc, ctx, env);
            std::vector<double> stress(5 * (size_t)n), fnet(3 * (size_t)n);
            if ((rc = bcg_mean_stress(ctx, stress.data(), fnet.data()))) return fail("bcg_mean_stress", rc, ctx, env);
            if (!write_file(dir + "/out" + std::to_string(k) + "_stress.bin", stress.data(),
                            stress.size() * sizeof(double)))
                return 3;
            moved.resize(n);
        }
        int32_t nd = 0;
        rc = bcg_sim_advance(ctx, &clk, out ? moved.data() : nullptr, (int32_t)moved.size(), &nd, lin.data(), 4096);
        if (rc) return fail("bcg_sim_advance", rc, ctx, env);
        if (out) {
            if (!write_file(dir + "/out" + std::to_string(k) + "_moved.bin", moved.data(),
                            moved.size() * sizeof(bcg_cell)))
                return 3;
            out_k++;
        }
        lineage.insert(lineage.end(), lin.begin(), lin.begin() + 2 * nd);
    }
    if ((rc = bcg_num_cells(ctx, &n))) return fail("bcg_num_cells", rc, ctx, env);
    cells.resize(n);
    int32_t got = 0;
    if ((rc = bcg_get_cells(ctx, cells.data(), n, &got))) return fail("bcg_get_cells", rc, ctx, env);
    int32_t draws[4];
    for (int i = 0; i < 4; i++)
        if ((rc = bcg_rand(ctx, &draws[i]))) return fail("bcg_rand", rc, ctx, env);
    if (!write_file(dir + "/lineage.bin", lineage.data(), lineage.size() * sizeof(int32_t)) ||
        !write_file(dir + "/cells_final.bin", cells.data(), cells.size() * sizeof(bcg_cell)) ||
        !write_file(dir + "/draws.bin", draws, sizeof draws))
        return 3;
    bcg_env_destroy(env);
    bcg_destroy(ctx);
    std::printf("loop: %d steps, %d output steps, %d cells, %zu divisions\n", steps, out_k, (int)n,
                lineage.size() / 2);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc == 2 && std::string(argv[1]) == "abi") return abi();
    if (argc == 3 && std::string(argv[1]) == "loop") return loop(argv[2]);
    std::fprintf(stderr, "usage: %s abi | loop <dir>\n", argv[0]);
    return 2;
}
