#ifndef _B2R_CONFIG_
#define _B2R_CONFIG_
/* test configuration in the form of the reference's b2r_config.h:6-17:
 * a 48 x 40 x 32 environment, B2R_D 2 as there, frozen faces, 3D */
#define ENV_DIM_X           48
#define ENV_DIM_Y           40
#define ENV_DIM_Z           32
#define GBL_ENV             ((ENV_DIM_X)*(ENV_DIM_Y)*(ENV_DIM_Z))
#define B2R_D               2
#define _PERIODIC_          0
#define _ENV_3D_            1
#define _DEBUG_             0
#define _MDL_DEV_           1
typedef float CELL_DT;
#endif
