/*
 * Test fixture: definitions of the reference's scalar globals that the hot-path
 * shim writes (names and layouts of include/cmc/cmc_vars.h:11-12,32-33,38,188,293
 * and of clus_struct_t / Etotal_struct_t, include/cmc/cmc.h:1690-1745).  Loaded
 * with RTLD_GLOBAL before libcmchost.so, it stands where the reference executable's
 * own definitions would, so the shim's weak references bind to these.
 */
typedef struct { long N_MAX, N_MAX_NEW, N_STAR, N_STAR_NEW, N_BINARY; } clus_struct_t;
typedef struct { double tot, New, ini, K, P, Eint, Eb; } Etotal_struct_t;

clus_struct_t clus;
Etotal_struct_t Etotal;
double Mtotal, Rtidal, max_r, Dt, TotalTime, TidalMassLoss, Eescaped, Jescaped, Etidal, Eoops;
