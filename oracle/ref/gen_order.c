/*
 * gen_order.c -- writes the reference's own block numbering: the level-0
 * Z-order that AMR_set_coord (AMR.c:168-275) of tp2872/HAMR_gamma_wind assigns
 * to the blocks of a single-level grid of NB_1 x NB_2 x NB_3 blocks (the
 * numbers the rdump<n> restart files carry, restart.c:41-42).
 *
 * Test infrastructure: oracle/ref/Makefile compiles AMR.c where it lies, once
 * per block count (the counts are compile-time macros of config.h), and links
 * it with this file; the output is the committed fixture
 * tests/golden/ref_block_order.json.  Prints one JSON object: the block
 * counts and, for n = 0 .. n_max, block n's coordinates (AMR_COORD1..3).
 */
#include "include.h"
#include "decs_MPI.h"
#include "defs.h"

void AMR_set_coord(void);

int main(void)
{
    int n, l;
    block = calloc((size_t)NB_1 * NB_2 * NB_3 + 8, sizeof *block);
    for (l = 0; l < N_LEVELS; l++) {
        lin_coord[l] = calloc((size_t)NB_1 * NB_2 * NB_3 + 8, sizeof(int));
        lin_coord_RM[l] = calloc((size_t)NB_1 * NB_2 * NB_3 + 8, sizeof(int));
    }
    AMR_set_coord();
    printf("{\"nb\": [%d, %d, %d], \"levels\": %d, \"n_max\": %d, \"coords\": [", NB_1, NB_2, NB_3, N_LEVELS, n_max);
    for (n = 0; n <= n_max; n++)
        printf("%s[%d, %d, %d, %d]", n ? ", " : "", block[n][AMR_LEVEL], block[n][AMR_COORD1], block[n][AMR_COORD2],
               block[n][AMR_COORD3]);
    printf("]}\n");
    return 0;
}
