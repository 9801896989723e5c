/* dgesv_c.c -- a C caller of libmagma_amd written the way a MAGMA application
 * is: magma_init, a queue, device buffers from magma_malloc, host -> device
 * copies with magma_setmatrix, pointer arrays built on the device with
 * magma_dset_pointer / magma_ddisplace_pointers / magma_iset_pointer, the
 * batched solve, device -> host copies with magma_getmatrix / magma_getvector,
 * then the teardown.  (Reference runtime: interface_cuda/interface.cpp:690-716,
 * interface_cuda/copy_v2.cpp; pointer helpers magmablas/zset_pointer.cu:87-288,
 * magmablas/set_pointer.cu:59-119.)
 *
 *   dgesv_c IN OUT
 * IN : int32 count, n, nrhs; then count*n*n doubles (A, column-major, ld n),
 *      count*n*nrhs doubles (B, ld n)
 * OUT: int32 return code; count ints info; count*n ints ipiv; count*n*n
 *      doubles LU; count*n*nrhs doubles X
 *
 * The matrices are stored on the device with a padded leading dimension
 * ldda = n + 3, each starting one row down (row offset 1), to exercise the
 * pointer helpers' row/column/stride arguments. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "magma_v2.h"

static int fail(const char *what, long code)
{
    fprintf(stderr, "dgesv_c: %s failed (%ld)\n", what, code);
    return 2;
}

int main(int argc, char **argv)
{
    if (argc != 3) {
        fprintf(stderr, "usage: dgesv_c IN OUT\n");
        return 2;
    }
    FILE *fi = fopen(argv[1], "rb");
    if (!fi) return fail("open input", 0);
    int hdr[3];
    if (fread(hdr, sizeof(int), 3, fi) != 3) return fail("read header", 0);
    const magma_int_t count = hdr[0], n = hdr[1], nrhs = hdr[2];
    const magma_int_t ldda = n + 3, lddb = n + 1;
    const size_t na = (size_t)count * n * n, nb = (size_t)count * n * nrhs;

    double *hA = NULL, *hB = NULL;
    magma_int_t *hinfo = NULL, *hipiv = NULL;
    if (magma_malloc_cpu((void **)&hA, na * sizeof(double)) != MAGMA_SUCCESS) return fail("magma_malloc_cpu", 0);
    if (magma_malloc_pinned((void **)&hB, nb * sizeof(double)) != MAGMA_SUCCESS) return fail("magma_malloc_pinned", 0);
    if (magma_malloc_cpu((void **)&hinfo, count * sizeof(magma_int_t)) != MAGMA_SUCCESS ||
        magma_malloc_cpu((void **)&hipiv, (size_t)count * n * sizeof(magma_int_t)) != MAGMA_SUCCESS)
        return fail("magma_malloc_cpu", 0);
    if (fread(hA, sizeof(double), na, fi) != na || fread(hB, sizeof(double), nb, fi) != nb)
        return fail("read matrices", 0);
    fclose(fi);

    magma_int_t ret = magma_init();
    if (ret != MAGMA_SUCCESS) return fail("magma_init", ret);
    magma_device_t dev;
    magma_getdevice(&dev);
    magma_queue_t queue = NULL;
    magma_queue_create(dev, &queue);
    if (!queue) return fail("magma_queue_create", 0);
    if (magma_queue_get_device(queue) != dev) return fail("magma_queue_get_device", magma_queue_get_device(queue));

    /* one buffer for all matrices: (ldda x n) per matrix, plus a spare row */
    double *dA = NULL, *dB = NULL;
    magma_int_t *dipiv = NULL, *dinfo = NULL;
    double **dA0_array = NULL, **dA_array = NULL, **dB_array = NULL;
    magma_int_t **dipiv_array = NULL;
    if (magma_dmalloc(&dA, (size_t)ldda * n * count + 1) != MAGMA_SUCCESS ||
        magma_dmalloc(&dB, (size_t)lddb * nrhs * count) != MAGMA_SUCCESS ||
        magma_imalloc(&dipiv, (size_t)n * count) != MAGMA_SUCCESS ||
        magma_imalloc(&dinfo, count) != MAGMA_SUCCESS ||
        magma_malloc((magma_ptr *)&dA0_array, count * sizeof(double *)) != MAGMA_SUCCESS ||
        magma_malloc((magma_ptr *)&dA_array, count * sizeof(double *)) != MAGMA_SUCCESS ||
        magma_malloc((magma_ptr *)&dB_array, count * sizeof(double *)) != MAGMA_SUCCESS ||
        magma_malloc((magma_ptr *)&dipiv_array, count * sizeof(magma_int_t *)) != MAGMA_SUCCESS)
        return fail("magma_malloc", 0);

    /* the whole batch as one ldda x (n count) matrix, rows 1..n of each column */
    magma_setmatrix(n, n * count, sizeof(double), hA, n, dA + 1, ldda, queue);
    magma_setmatrix(n, nrhs * count, sizeof(double), hB, n, dB, lddb, queue);

    /* dA0_array[i] = dA + i ldda n (row 0), then displaced one row down */
    magma_dset_pointer(dA0_array, dA, ldda, 0, 0, ldda * n, count, queue);
    magma_ddisplace_pointers(dA_array, dA0_array, ldda, 1, 0, count, queue);
    magma_dset_pointer(dB_array, dB, lddb, 0, 0, lddb * nrhs, count, queue);
    magma_iset_pointer(dipiv_array, dipiv, 1, 0, 0, n, count, queue);

    ret = magma_dgesv_batched(n, nrhs, dA_array, ldda, dipiv_array, dB_array, lddb, dinfo, count, queue);
    if (ret != MAGMA_SUCCESS) fprintf(stderr, "dgesv_c: magma_dgesv_batched returned %d (%s)\n", (int)ret,
                                      magma_strerror(ret));

    magma_getvector(count, sizeof(magma_int_t), dinfo, 1, hinfo, 1, queue);
    magma_getvector(n * count, sizeof(magma_int_t), dipiv, 1, hipiv, 1, queue);
    magma_getmatrix(n, n * count, sizeof(double), dA + 1, ldda, hA, n, queue);
    magma_getmatrix(n, nrhs * count, sizeof(double), dB, lddb, hB, n, queue);
    magma_queue_sync(queue);

    FILE *fo = fopen(argv[2], "wb");
    if (!fo) return fail("open output", 0);
    int r32 = (int)ret;
    fwrite(&r32, sizeof(int), 1, fo);
    fwrite(hinfo, sizeof(magma_int_t), count, fo);
    fwrite(hipiv, sizeof(magma_int_t), (size_t)n * count, fo);
    fwrite(hA, sizeof(double), na, fo);
    fwrite(hB, sizeof(double), nb, fo);
    fclose(fo);

    magma_free(dA);
    magma_free(dB);
    magma_free(dipiv);
    magma_free(dinfo);
    magma_free(dA0_array);
    magma_free(dA_array);
    magma_free(dB_array);
    magma_free(dipiv_array);
    magma_queue_destroy(queue);
    magma_free_cpu(hA);
    magma_free_pinned(hB);
    magma_free_cpu(hinfo);
    magma_free_cpu(hipiv);
    magma_finalize();
    return ret == MAGMA_SUCCESS ? 0 : 1;
}
