/* Prints the layout of the C ABI's structs as the C compiler lays them out,
 * for the ctypes mirror check in tests/test_capi_host.py. */
#include <stddef.h>
#include <stdio.h>

#include "etn/etn.h"

#define FIELD(T, f) printf("%s %s %zu\n", #T, #f, offsetof(T, f))

int main(void)
{
    printf("etn_reconstruction_desc sizeof %zu\n", sizeof(etn_reconstruction_desc));
    FIELD(etn_reconstruction_desc, algorithm);
    FIELD(etn_reconstruction_desc, block_size);
    FIELD(etn_reconstruction_desc, iterations);
    FIELD(etn_reconstruction_desc, lambda);
    FIELD(etn_reconstruction_desc, use_long_object_compensation);
    FIELD(etn_reconstruction_desc, fp_samples);
    FIELD(etn_reconstruction_desc, bp_samples);
    FIELD(etn_reconstruction_desc, projection_iteration);
    FIELD(etn_reconstruction_desc, skip_projections);
    FIELD(etn_reconstruction_desc, skip_count);
    FIELD(etn_reconstruction_desc, voxel_type);
    FIELD(etn_reconstruction_desc, volume);
    FIELD(etn_reconstruction_desc, init_value);
    FIELD(etn_reconstruction_desc, base_geometry);
    FIELD(etn_reconstruction_desc, projection_count);
    FIELD(etn_reconstruction_desc, tilt_angles);
    FIELD(etn_reconstruction_desc, x_axis_tilt);
    FIELD(etn_reconstruction_desc, projections);
    FIELD(etn_reconstruction_desc, mask);
    FIELD(etn_reconstruction_desc, flags);
    FIELD(etn_reconstruction_desc, geometries);
    printf("etn_scanner_geometry sizeof %zu\n", sizeof(etn_scanner_geometry));
    printf("etn_volume_geometry sizeof %zu\n", sizeof(etn_volume_geometry));
    return 0;
}
