// b2r_output.cpp -- the reference's on-disk outputs, same file names and
// layouts:
//   write_block_data_bin  eqwp.cu:616-627  displacement-<ts>.bin: three int32
//                         (Ngx, Ngy, Ngz) then Ngx*Ngy*Ngz FP32, x fastest;
//   write_block_data      eqwp.cu:636-650  displacement-<ts>.txt: one line per
//                         (z, y) row, every value followed by ',';
//   print_block_data_step heat-3D.hpp:105-127  R-<R>-rank-<rank>-i-<step>.txt:
//                         per z a "slice: [z, :, :]" line, then one line per
//                         row of "%0.2f," values over the padded block.
// Files go to the working directory, as in the reference.  Errors print the
// reason and exit(-1), like the other shims.
#include <cstdio>
#include <cstdlib>
#include <fstream>

#include "b2r_host.h"

int B2R_OUTPUT = 0;

static void fail(const char* name)
{
    fprintf(stderr, "B2R Error: cannot write %s\n", name);
    exit(-1);
}

void write_block_data_bin(int time_step, CELL_DT* p_data, int Ngx, int Ngy, int Ngz)
{
    char name[100];
    snprintf(name, sizeof(name), "displacement-%d.bin", time_step);
    std::ofstream f(name, std::ios::binary);
    const int header[3] = {Ngx, Ngy, Ngz};
    f.write((const char*)header, sizeof(header));
    f.write((const char*)p_data, (std::streamsize)sizeof(CELL_DT) * Ngx * Ngy * Ngz);
    if (!f) fail(name);
}

void write_block_data(int time_step, CELL_DT* p_data, int Ngx, int Ngy, int Ngz)
{
    char name[100];
    snprintf(name, sizeof(name), "displacement-%d.txt", time_step);
    std::ofstream f(name);
    for (long long zy = 0; zy < (long long)Ngz * Ngy; ++zy) {
        const CELL_DT* row = p_data + zy * Ngx;
        for (int x = 0; x < Ngx; ++x) f << row[x] << ",";  // default stream format
        f << "\n";
    }
    if (!f) fail(name);
}

void print_block_data_step(int rank, float step, const CELL_DT* block)
{
    char name[100];
    snprintf(name, sizeof(name), "R-%d-rank-%d-i-%0.1f.txt", B2R_R, rank, step);
    FILE* f = fopen(name, "w");
    if (!f) fail(name);
    const long long sx = B2R_BLOCK_SIZE_X, sy = B2R_BLOCK_SIZE_Y;
    for (int z = 0; z < B2R_BLOCK_SIZE_Z; ++z) {
        fprintf(f, "slice: [%d, :, :]\n", z);
        for (int y = 0; y < B2R_BLOCK_SIZE_Y; ++y) {
            const CELL_DT* row = block + (z * sy + y) * sx;
            for (int x = 0; x < B2R_BLOCK_SIZE_X; ++x) fprintf(f, "%0.2f,", row[x]);
            fputc('\n', f);
        }
    }
    if (fclose(f)) fail(name);
}
