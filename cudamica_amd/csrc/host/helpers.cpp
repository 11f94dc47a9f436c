/*
 * helpers.cpp -- sample partition and host-side matrix file I/O
 * (fraimondo/cudamica src/helpers.cu:19-49, 513-630).
 *
 * Formats: data files (.fdt) are raw float32, channels fastest
 * (x[t*nchannels + i]); result files are raw float64 rows of xsize
 * elements.  Files are mmapped; a file shorter than the requested matrix is
 * an ERRORIO (the reference reads past the end), and open() failures are
 * detected with fd < 0 (the reference tests fd == 0).
 */
#include "../../../include/helpers.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <cstring>

#include "../../../include/config.h"

namespace {

/* Maps filename read-only; returns nullptr with *err set on failure. */
const void* map_file(const char* filename, size_t need, size_t* size, int* fd_out, error* err) {
    int fd = open(filename, O_RDONLY);
    if (fd < 0) { *err = ERRORNOFILE; return nullptr; }
    struct stat sb;
    if (fstat(fd, &sb) == -1) {
        close(fd);
        *err = ERRORNOFILE;
        return nullptr;
    }
    if ((size_t)sb.st_size < need) {
        fprintf(stderr, "Data file %s holds %ld bytes, %zu needed\n", filename, (long)sb.st_size, need);
        close(fd);
        *err = ERRORIO;
        return nullptr;
    }
    void* p = mmap(nullptr, need ? need : 1, PROT_READ, MAP_SHARED, fd, 0);
    if (p == MAP_FAILED) {
        close(fd);
        *err = ERRORNOFILE;
        return nullptr;
    }
    *size = need ? need : 1;
    *fd_out = fd;
    *err = SUCCESS;
    return p;
}

void unmap_file(const void* p, size_t size, int fd) {
    munmap((void*)p, size);
    close(fd);
}

}  // namespace

extern "C" {

error recalc_config(void) {
    natural nsamples = currentConfig.nsamples;
    natural maxblocksize = currentConfig.block_size;
    natural ndevices = currentConfig.devcount ? currentConfig.devcount : 1;
    if (maxblocksize == 0 || nsamples == 0) return ERRORINVALIDPARAM;
    if (maxblocksize > nsamples) maxblocksize = nsamples;
    /* The reference shrinks the block size until it divides the samples
     * (helpers.cu:28-40).  The engine handles a ragged last block, so the
     * block size is kept and the last block is shorter. */
    natural blockgpumodel = (nsamples + maxblocksize * ndevices - 1) / (maxblocksize * ndevices);
    currentConfig.block_size = maxblocksize;
    currentConfig.blocks_per_gpu_per_model = blockgpumodel;
    return SUCCESS;
}

error loadMatrix(real* data, natural rows, natural elems, char* filename, bool isdouble) {
    const size_t cnt = (size_t)rows * elems;
    const size_t need = cnt * (isdouble ? sizeof(double) : sizeof(float));
    size_t sz = 0;
    int fd = -1;
    error err = SUCCESS;
    const void* p = map_file(filename, need, &sz, &fd, &err);
    if (!p) return err;
    if (isdouble) {
        std::memcpy(data, p, need);
    } else {
        const float* f = (const float*)p;
        for (size_t i = 0; i < cnt; ++i) data[i] = (real)f[i];
    }
    unmap_file(p, sz, fd);
    return SUCCESS;
}

error loadMatrixF32(float* data, natural rows, natural elems, const char* filename) {
    const size_t need = (size_t)rows * elems * sizeof(float);
    size_t sz = 0;
    int fd = -1;
    error err = SUCCESS;
    const void* p = map_file(filename, need, &sz, &fd, &err);
    if (!p) return err;
    std::memcpy(data, p, need);
    unmap_file(p, sz, fd);
    return SUCCESS;
}

error writeValue(real value, char* filename) {
    FILE* out = fopen(filename, "w");
    if (!out) return ERRORNOFILE;
    error r = fwrite(&value, sizeof(real), 1, out) == 1 ? SUCCESS : ERRORIO;
    fclose(out);
    return r;
}

error writeMatrix(amica_pitched ptr, char* filename) {
    FILE* out = fopen(filename, "w");
    if (!out) return ERRORNOFILE;
    const char* base = (const char*)ptr.ptr;
    error r = SUCCESS;
    for (size_t i = 0; i < ptr.ysize; ++i) {
        if (fwrite(base + i * ptr.pitch, sizeof(real), ptr.xsize, out) != ptr.xsize) {
            r = ERRORIO;
            break;
        }
    }
    fclose(out);
    return r;
}

}  // extern "C"
