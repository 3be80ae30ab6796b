// ref_shim.cpp -- C-linkage entry points into the reference's CPU heat path,
// linked with /root/reference/heat-3D.cpp into oracle/_ref/libheat3d_ref.so.
// TEST INFRASTRUCTURE ONLY (the checker and the bench's CPU baseline).
// The declarations restate the signatures of heat-3D.cpp:48, 64, 85, 109, 136.
typedef float CELL_DT;
void cpu_fdm_heat_diffuse1(float* h_temp1, float* h_temp2, int Ngx, int Ngy, int Ngz, int b2r_i);
void cpu_fdm_heat_diffuse2(float* h_temp1, float* h_temp2, int Ngx, int Ngy, int Ngz, int b2r_i);
float heat_3D_cpu_main(CELL_DT* h_in, int b2r_R, int b2r_D, int Ngx, int Ngy, int Ngz);

extern "C" void ref_cpu_fdm_heat_diffuse1(float* a, float* b, int nx, int ny, int nz, int i)
{
    cpu_fdm_heat_diffuse1(a, b, nx, ny, nz, i);
}

extern "C" float ref_heat_3D_cpu_main(float* h, int R, int D, int nx, int ny, int nz)
{
    return heat_3D_cpu_main(h, R, D, nx, ny, nz);
}
