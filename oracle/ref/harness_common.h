// Dump helpers for the reference host-code harnesses (test infrastructure).
#pragma once
#include <cstdio>
#include <cstdlib>
#include <string>

static std::string g_outdir;

template <class T> static void dump(const char *name, const T *p, long n) {
    std::string path = g_outdir + "/" + name;
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) {
        fprintf(stderr, "cannot write %s\n", path.c_str());
        exit(3);
    }
    fwrite(p, sizeof(T), (size_t)n, f);
    fclose(f);
}

static void init_outdir() {
    const char *o = getenv("FD23_REF_OUT");
    if (!o) {
        fprintf(stderr, "set FD23_REF_OUT to the output directory\n");
        exit(2);
    }
    g_outdir = o;
}
