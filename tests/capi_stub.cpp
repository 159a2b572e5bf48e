// A C++ consumer of include/athad_hip.h, shaped like the hooks a maintainer adds to
// cAtmosphereModel (cAtmosphereModel.cpp:667 set-up, :1103-1647 loop): it holds its state in
// the reference's Array layout (double*** x, one separately allocated km-row per (i, j),
// lib/Array.cpp:62-71) and moves it through athad_upload_rows / athad_download_rows.
//
//   capi_stub null                          calls on a null context and bad arguments
//   capi_stub run im jm km steps out.bin    create (FP64, lex), init, download every field into
//                                           row storage and upload it back, set_iter, steps,
//                                           download the 35 fields, write them to out.bin
//                                           ([field][i][j][k] doubles), destroy
// Prints one JSON line of status codes.
#include "athad_hip.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

// The reference Array: rows allocated one by one, in reverse order, so consecutive (i, j)
// rows are not adjacent in memory.
struct RowArray {
    int im, jm, km;
    std::vector<double*> rows;
    RowArray(int im_, int jm_, int km_) : im(im_), jm(jm_), km(km_), rows((size_t)im_ * jm_) {
        for (size_t r = rows.size(); r-- > 0;) rows[r] = (double*)std::calloc((size_t)km + 3, sizeof(double)) + 1;
    }
    ~RowArray() {
        for (double* p : rows) std::free(p - 1);
    }
};

int null_calls() {
    athad_grid g{11, 21, 41};
    athad_ctx* ctx = nullptr;
    double* rows[1] = {nullptr};
    int it = 0, tot = 0;
    const int bad_create = athad_create(0, &g, nullptr, ATHAD_FP64, ATHAD_ORDER_LEX, &ctx);
    const int up = athad_upload_rows(nullptr, 0, ATHAD_T, rows);
    const int down = athad_download_rows(nullptr, 0, ATHAD_T, rows);
    const int si = athad_set_iter(nullptr, 0, 1, 1);
    const int gi = athad_get_iter(nullptr, 0, &it, &tot);
    const int st = athad_step(nullptr);
    const int rk = athad_rk4(nullptr);
    const int tm = athad_set_turbulence_model(nullptr, ATHAD_TURB_K_EPSILON);
    const int ti = athad_turbulence_init(nullptr);
    const int ds = athad_destroy(nullptr);
    const char* msg = athad_last_error(nullptr);
    const long long fp = (long long)athad_footprint_bytes(&g, ATHAD_FP64);
    const long long rb = (long long)athad_restart_bytes(&g);
    const double dew = athad_dew_point(60000.0);
    athad_member_params p{};
    const int mfs = athad_member_from_surface(60.0, 513.15, 0.01, 0.92, &p);
    std::printf("{\"create_bad\": %d, \"upload_rows\": %d, \"download_rows\": %d, \"set_iter\": %d, "
                "\"get_iter\": %d, \"step\": %d, \"rk4\": %d, \"turb_model\": %d, \"turb_init\": %d, "
                "\"destroy\": %d, \"last_error\": \"%s\", "
                "\"footprint\": %lld, \"restart_bytes\": %lld, \"dew_point_60bar\": %.6f, "
                "\"member_from_surface\": %d, \"x_H2O\": %.6f, \"ctx_null\": %d, \"version\": \"%s\"}\n",
                bad_create, up, down, si, gi, st, rk, tm, ti, ds, msg ? msg : "", fp, rb, dew, mfs, p.x_H2O,
                ctx == nullptr, athad_version());
    return 0;
}

int run(int im, int jm, int km, int steps, const char* out) {
    athad_grid g{im, jm, km};
    athad_ctx* ctx = nullptr;
    int s = athad_create(1, &g, nullptr, ATHAD_FP64, ATHAD_ORDER_LEX, &ctx);
    if (s != ATHAD_OK) {
        std::printf("{\"create\": %d, \"error\": \"%s\"}\n", s, athad_last_error(nullptr));
        return 1;
    }
    int worst = ATHAD_OK;
    auto chk = [&](int r) { if (r != ATHAD_OK && worst == ATHAD_OK) worst = r; };
    chk(athad_init_state(ctx, 1234));
    // round trip of every field through the reference's row layout
    std::vector<RowArray*> arrays;
    for (int f = 0; f < ATHAD_NFIELD; f++) {
        arrays.push_back(new RowArray(im, jm, km));
        chk(athad_download_rows(ctx, 0, f, arrays.back()->rows.data()));
    }
    for (int f = 0; f < ATHAD_NFIELD; f++) chk(athad_upload_rows(ctx, 0, f, arrays[f]->rows.data()));
    chk(athad_set_iter(ctx, 0, 0, 0));
    for (int n = 0; n < steps; n++) chk(athad_step(ctx));
    int it = -1, tot = -1;
    chk(athad_get_iter(ctx, 0, &it, &tot));
    FILE* fo = std::fopen(out, "wb");
    if (!fo) return 2;
    for (int f = 0; f < ATHAD_NFIELD; f++) {
        chk(athad_download_rows(ctx, 0, f, arrays[f]->rows.data()));
        for (double* row : arrays[f]->rows) std::fwrite(row, sizeof(double), (size_t)km, fo);
    }
    std::fclose(fo);
    // a null row pointer is refused
    double* keep = arrays[0]->rows[3];
    arrays[0]->rows[3] = nullptr;
    const int null_row = athad_download_rows(ctx, 0, 0, arrays[0]->rows.data());
    arrays[0]->rows[3] = keep;
    for (RowArray* a : arrays) delete a;
    const int d = athad_destroy(ctx);
    std::printf("{\"status\": %d, \"iter_n\": %d, \"total\": %d, \"null_row\": %d, \"destroy\": %d}\n", worst, it, tot,
                null_row, d);
    return worst == ATHAD_OK ? 0 : 3;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc >= 2 && std::strcmp(argv[1], "null") == 0) return null_calls();
    if (argc >= 7 && std::strcmp(argv[1], "run") == 0)
        return run(std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]), argv[6]);
    std::fprintf(stderr, "usage: %s null | run im jm km steps out.bin\n", argv[0]);
    return 64;
}
