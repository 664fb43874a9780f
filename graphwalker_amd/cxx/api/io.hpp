// api/io.hpp -- whole-buffer positional reads and writes (the role of
// preada / pwritea in src/api/io.hpp:11-76), retrying short transfers.
#ifndef GW_CXX_IO_HPP
#define GW_CXX_IO_HPP

#include <assert.h>
#include <errno.h>
#include <string.h>
#include <unistd.h>

#include <cstdint>

#include "logger/logger.hpp"

template <typename T>
void preada(int f, T* tbuf, size_t nbytes, size_t off) {
    char* buf = reinterpret_cast<char*>(tbuf);
    size_t done = 0;
    while (done < nbytes) {
        ssize_t r = pread(f, buf + done, nbytes - done, static_cast<off_t>(off + done));
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) {
            logstream(LOG_FATAL) << "preada: read failed at offset " << off + done << ": " << strerror(errno) << std::endl;
        }
        done += static_cast<size_t>(r);
    }
}

template <typename T>
void pwritea(int f, const T* tbuf, size_t nbytes, size_t off = 0) {
    const char* buf = reinterpret_cast<const char*>(tbuf);
    size_t done = 0;
    while (done < nbytes) {
        ssize_t r = pwrite(f, buf + done, nbytes - done, static_cast<off_t>(off + done));
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) {
            logstream(LOG_FATAL) << "pwritea: write failed at offset " << off + done << ": " << strerror(errno) << std::endl;
        }
        done += static_cast<size_t>(r);
    }
}

#endif
