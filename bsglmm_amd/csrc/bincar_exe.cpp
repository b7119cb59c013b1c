// BinCar: the reference program's command line (bsg_bincar_main in libbsglmm.so).
#include "../../include/bsglmm_io.h"

int main(int argc, char **argv) { return bsg_bincar_main(argc, argv); }
