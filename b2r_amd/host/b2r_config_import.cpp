// b2r_config_import.cpp -- the compile-time switches of a b2r_config.h
// (the reference's b2r_config.h:6-13), read in a translation unit of their own
// because the header's macros carry the names of the run-time globals that
// b2r_host.cpp defines.  make -C b2r_amd/host CONFIG=/path/to/b2r_config.h
// compiles this file with -DB2R_CONFIG_HEADER; without it the defaults below
// stand in (a 256^3 3D environment, B2R_D 1, no periodic faces).
#ifdef B2R_CONFIG_HEADER
#include B2R_CONFIG_HEADER
extern const int b2r_cfg_env[3] = {ENV_DIM_X, ENV_DIM_Y, ENV_DIM_Z};
extern const int b2r_cfg_d = B2R_D;
extern const int b2r_cfg_periodic = _PERIODIC_;
extern const int b2r_cfg_env_3d = _ENV_3D_;
#else
extern const int b2r_cfg_env[3] = {256, 256, 256};
extern const int b2r_cfg_d = 1;
extern const int b2r_cfg_periodic = 0;
extern const int b2r_cfg_env_3d = 1;
#endif
