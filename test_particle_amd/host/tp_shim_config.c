/*
 * tp_shim_config.c -- reads the optional "GPU tracking" section of init.dat
 * (layout and defaults in tp_shim_config.h).
 *
 * The section is found the way the reference finds its own sections, a
 * strstr over whole lines (src/domain.c:86-90); its lines are "Label: value"
 * like the reference's fixed fscanf lines (src/domain.c:92-96), but each is
 * optional, may come in any order and is matched by the start of its label.
 */
#include <ctype.h>
#include <errno.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "tp_gpu.h"
#include "tp_shim_config.h"

#define TP_CFG_LINE 1024
#define TP_CFG_ANCHOR "GPU tracking"

void tp_shim_config_defaults(tp_shim_config *c)
{
    c->found = 0;
    c->pusher = TP_PUSH_WIRZ;
    c->precision = TP_FP64;
    c->flags = 0;
    c->bs_rtol = 1.0e-14;
    c->bs_atol = 1.0e-5;
    c->bs_h0 = 1.0e-5;
    c->bs_kmax = 8;
}

static int cfg_fail(char *err, size_t nerr, int line, const char *fmt, ...)
{
    if (err && nerr > 0) {
        int k = line > 0 ? snprintf(err, nerr, "init.dat line %d: ", line)
                         : snprintf(err, nerr, "init.dat: ");
        if (k < 0) k = 0;
        if ((size_t)k < nerr) {
            va_list ap;
            va_start(ap, fmt);
            vsnprintf(err + k, nerr - (size_t)k, fmt, ap);
            va_end(ap);
        }
    }
    return -1;
}

/* the value after the line's last ':' with surrounding blanks removed, in
 * place; NULL when the line has no ':' */
static char *cfg_value(char *line)
{
    char *colon = strrchr(line, ':');
    if (!colon) return NULL;
    char *v = colon + 1;
    while (*v && isspace((unsigned char)*v)) v++;
    char *end = v + strlen(v);
    while (end > v && isspace((unsigned char)end[-1])) *--end = '\0';
    return v;
}

static int cfg_double(const char *v, double lo, double hi, double *out)
{
    char *end;
    errno = 0;
    double x = strtod(v, &end);
    if (end == v || *end != '\0' || errno == ERANGE || !isfinite(x) || x < lo || x > hi)
        return -1;
    *out = x;
    return 0;
}

static int cfg_int(const char *v, long lo, long hi, int *out)
{
    char *end;
    errno = 0;
    long x = strtol(v, &end, 10);
    if (end == v || *end != '\0' || errno == ERANGE || x < lo || x > hi) return -1;
    *out = (int)x;
    return 0;
}

/* one line of the section; 0 when it was understood */
static int cfg_line(char *line, int lineno, tp_shim_config *c, char *err, size_t nerr)
{
    char *label = line;
    while (*label && isspace((unsigned char)*label)) label++;
    char *v = cfg_value(label);
    if (!v) return cfg_fail(err, nerr, lineno, "expected \"Label: value\" in the GPU tracking "
                            "section");
    int flag;
    if (!strncmp(label, "Pusher", 6)) {
        if (!strcmp(v, "wirz")) c->pusher = TP_PUSH_WIRZ;
        else if (!strcmp(v, "boris")) c->pusher = TP_PUSH_BORIS;
        else return cfg_fail(err, nerr, lineno, "pusher \"%s\" is neither wirz nor boris", v);
    } else if (!strncmp(label, "Precision", 9)) {
        if (!strcmp(v, "fp64")) c->precision = TP_FP64;
        else if (!strcmp(v, "fp32")) c->precision = TP_FP32_MIXED;
        else return cfg_fail(err, nerr, lineno, "precision \"%s\" is neither fp64 nor fp32", v);
    } else if (!strncmp(label, "Check for non-finite", 20)) {
        if (cfg_int(v, 0, 1, &flag)) return cfg_fail(err, nerr, lineno, "expected 0 or 1");
        c->flags = flag ? c->flags | TP_FLAG_CHECK_FINITE : c->flags & ~TP_FLAG_CHECK_FINITE;
    } else if (!strncmp(label, "Fix stale diffusion", 19)) {
        if (cfg_int(v, 0, 1, &flag)) return cfg_fail(err, nerr, lineno, "expected 0 or 1");
        c->flags = flag ? c->flags | TP_FLAG_FIX_STALE_DCOEFF
                        : c->flags & ~TP_FLAG_FIX_STALE_DCOEFF;
    } else if (!strncmp(label, "Adaptive relative tolerance", 27)) {
        if (cfg_double(v, 0.0, 1.0, &c->bs_rtol))
            return cfg_fail(err, nerr, lineno, "relative tolerance \"%s\" is not in [0, 1]", v);
    } else if (!strncmp(label, "Adaptive absolute tolerance", 27)) {
        if (cfg_double(v, 0.0, 1.0e300, &c->bs_atol))
            return cfg_fail(err, nerr, lineno, "absolute tolerance \"%s\" is not a finite "
                            "number >= 0", v);
    } else if (!strncmp(label, "Adaptive first step", 19)) {
        if (cfg_double(v, 1.0e-300, 1.0e300, &c->bs_h0))
            return cfg_fail(err, nerr, lineno, "first step \"%s\" is not a finite number > 0",
                            v);
    } else if (!strncmp(label, "Adaptive maximum sequence index", 31)) {
        if (cfg_int(v, 3, 8, &c->bs_kmax))
            return cfg_fail(err, nerr, lineno, "maximum sequence index \"%s\" is not in 3..8",
                            v);
    } else {
        return cfg_fail(err, nerr, lineno, "unknown line in the GPU tracking section");
    }
    return 0;
}

int tp_shim_config_read(const char *path, tp_shim_config *c, char *err, size_t nerr)
{
    tp_shim_config_defaults(c);
    if (err && nerr) err[0] = '\0';
    FILE *fp = fopen(path, "r");
    if (!fp) return 0;
    char line[TP_CFG_LINE];
    int lineno = 0, in = 0, rc = 0;
    while (fgets(line, sizeof(line), fp) != NULL) {
        lineno++;
        const size_t len = strlen(line);
        if (len == sizeof(line) - 1 && line[len - 1] != '\n' && !feof(fp)) {
            /* a line longer than the buffer: only an error inside the section */
            if (in) {
                rc = cfg_fail(err, nerr, lineno, "line longer than %d characters",
                              TP_CFG_LINE - 2);
                break;
            }
            int ch;
            while ((ch = fgetc(fp)) != EOF && ch != '\n') {}
            continue;
        }
        if (!in) {
            if (strstr(line, TP_CFG_ANCHOR) != NULL) {
                in = 1;
                c->found = 1;
            }
            continue;
        }
        const char *s = line;
        while (*s && isspace((unsigned char)*s)) s++;
        if (!strncmp(s, "---", 3)) break;       /* the section's closing rule */
        if (*s == '\0') continue;               /* blank line */
        if ((rc = cfg_line(line, lineno, c, err, nerr)) != 0) break;
    }
    fclose(fp);
    if (rc == 0 && c->precision == TP_FP32_MIXED && c->pusher != TP_PUSH_BORIS)
        rc = cfg_fail(err, nerr, 0, "precision fp32 needs pusher boris (the Wirz pusher "
                      "runs in FP64 only)");
    return rc;
}
