// Body of the fused 2D update kernels (flat2d.hip), included inside k_flat2d (x: the kernel
// argument) and k_flat2d_shots (x: the shot's row of the argument table); the template
// parameters RHO, PML, BETA, SAVE and the tile constants are those of flat2d.hip. The workgroup
// owns tile (blockIdx.x, blockIdx.y) of the grid's tiles.
    __shared__ float4 sU4[UY * UZ / 4 + 1];  // + a dump slot for threads without a chunk
    __shared__ float4 sB4[RHO == 2 ? BY * BZ / 4 + 1 : 1];
    __shared__ float sFy[RHO ? FY * kTZ : 1];
    __shared__ float sFz[RHO ? kTY * FZP : 1];
    float *sU = reinterpret_cast<float *>(sU4), *sB = reinterpret_cast<float *>(sB4);
    const int ny = g.n[1], nz = g.n[2];
    const int z0 = blockIdx.x * kTZ, y0 = blockIdx.y * kTY;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const long long bytes = g.total * 4;
    const rsrc_t ru = make(x.u, bytes);
    if (x.gout) {
        // sparse points owned by this tile, read from u in memory (no launch writes u)
        const int tile = blockIdx.y * gridDim.x + blockIdx.x;
        for (int e = x.goff[tile] + t; e < x.goff[tile + 1]; e += NT) {
            const int r = x.gpt[e];
            float acc = 0.f;
            for (int q = 0; q < x.gntap; ++q) {
                const long long pq = x.gidx[(long long)r * x.gntap + q];
                if (pq < 0) continue;
                acc += x.gw[(long long)r * x.gntap + q] * x.u[pq];
            }
            x.gout[r] = x.gsign * acc;
        }
    }

    // stage u (and b): rows past the padded plane read zero
    for (int e = t; e < UY * UC; e += NT) {
        const int row = e / UC, c = e - row * UC;
        const int y = y0 - 9 + row, z = z0 - 12 + 4 * c;
        const unsigned off = y <= ny + 9 ? (unsigned)((g.base + (long long)y * g.s1 + z) * 4) : kOOB;
        *reinterpret_cast<float4 *>(sU + row * UZ + 4 * c) = load4(ru, off);
    }
    if (RHO == 2) {
        const rsrc_t rb = make(x.b, bytes);
        for (int e = t; e < BY * BC; e += NT) {
            const int row = e / BC, c = e - row * BC;
            const int y = y0 - 5 + row, z = z0 - 8 + 4 * c;
            const unsigned off = y <= ny + 9 ? (unsigned)((g.base + (long long)y * g.s1 + z) * 4) : kOOB;
            *reinterpret_cast<float4 *>(sB + row * BZ + 4 * c) = load4(rb, off);
        }
    }
    lds_barrier();
    const int y = y0 + w, z = z0 + lane;
    const bool valid = y < ny && z < nz;
    const float *c0 = sU + (w + 9) * UZ + lane + 12;  // own cell in the u tile
    float K;
    if (RHO == 0) {
        float ky = g.w2[1][0] * c0[0], kz = g.w2[2][0] * c0[0];
#pragma unroll
        for (int r = 1; r <= 5; ++r) {
            ky += g.w2[1][r] * (c0[r * UZ] + c0[-r * UZ]);
            kz += g.w2[2][r] * (c0[r] + c0[-r]);
        }
        K = ky + kz;
    } else {
        // fluxes at half rows q = y0 - 5 + fr (axis 1) and half columns q = z0 - 5 + fz (axis 2)
        for (int e = t; e < FY * kTZ + kTY * FZN; e += NT) {
            if (e < FY * kTZ) {
                const int fr = e >> 6, l = e & 63;
                const float *c = sU + (fr + 4) * UZ + l + 12;
                float d = 0.f;
#pragma unroll
                for (int k = 0; k < 5; ++k) d += g.ws[1][k] * (c[(k + 1) * UZ] - c[-k * UZ]);
                if (RHO == 2) d *= 0.5f * (sB[fr * BZ + l + 8] + sB[(fr + 1) * BZ + l + 8]);
                sFy[fr * kTZ + l] = d;
            } else {
                const int ez = e - FY * kTZ;
                const int row = ez / FZN, fz = ez - row * FZN;
                const float *c = sU + (row + 9) * UZ + fz + 7;
                float d = 0.f;
#pragma unroll
                for (int k = 0; k < 5; ++k) d += g.ws[2][k] * (c[k + 1] - c[-k]);
                if (RHO == 2) d *= 0.5f * (sB[(row + 5) * BZ + fz + 3] + sB[(row + 5) * BZ + fz + 4]);
                sFz[row * FZP + fz] = d;
            }
        }
        lds_barrier();
        const float *fy = sFy + (w + 5) * kTZ + lane;
        const float *fz = sFz + w * FZP + lane + 5;
        float dy = 0.f, dz = 0.f;
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            dy += g.ws[1][k] * (fy[k * kTZ] - fy[-(k + 1) * kTZ]);
            dz += g.ws[2][k] * (fz[k] - fz[-(k + 1)]);
        }
        K = dy + dz;
    }
    if (!valid) return;
    const long long p = g.base + (long long)y * g.s1 + z;
    const float u = c0[0], o = x.other[p];
    float acc = x.G[p] * K;
    if (BETA) {
        float ly = g.w2[1][0] * u, lz = g.w2[2][0] * u;
#pragma unroll
        for (int r = 1; r <= 5; ++r) {
            ly += g.w2[1][r] * (c0[r * UZ] + c0[-r * UZ]);
            lz += g.w2[2][r] * (c0[r] + c0[-r]);
        }
        acc += x.Cbeta[p] * ((ly + lz) - x.lapother[p]);
        if (x.lapout) x.lapout[p] = ly + lz;  // the next step's Laplacian of its other level
    }
    if (PML) {
        float s1 = 0.f;
#pragma unroll
        for (int a = 1; a <= 2; ++a) {
            const int xa = a == 1 ? y : z;
            if (xa < g.nabs[a] || xa >= g.n[a] - g.nabs[a]) {
                const int st = a == 1 ? UZ : 1;
                float dF = 0.f, d2F = g.w2[a][0] * u;
#pragma unroll
                for (int r = 1; r <= 5; ++r) {
                    dF += g.w1[a][r] * (c0[r * st] - c0[-r * st]);
                    d2F += g.w2[a][r] * (c0[r * st] + c0[-r * st]);
                }
                const float s = x.pml[a][0][xa], al = x.pml[a][1][xa], ds = x.pml[a][2][xa], da = x.pml[a][3][xa];
                float *u1 = x.u1[a], *u2 = x.u2[a], *u3 = x.u3[a];
                const float dt = x.dt;
                float v3, v2, v1;
                if (!x.backward) {
                    v3 = u3[p] - dt * ((al + s) * u3[p] - s * s * (ds + da) * dF);
                    v2 = u2[p] - dt * ((al + s) * u2[p] - s * (2.f * ds * dF + da * dF + s * d2F) + v3);
                    v1 = u1[p] - dt * ((al + s) * u1[p] - (ds * dF + 2.f * s * d2F) + v2);
                } else {
                    v3 = u3[p] - dt * ((al + s) * u3[p] + s * s * (ds + da) * dF);
                    v2 = u2[p] - dt * ((al + s) * u2[p] + s * (2.f * ds * dF + da * dF) - s * s * d2F + v3);
                    v1 = u1[p] - dt * ((al + s) * u1[p] + ds * dF - 2.f * s * d2F + v2);
                }
                u3[p] = v3;
                u2[p] = v2;
                u1[p] = v1;
                s1 += v1;
            }
        }
        acc -= x.Cpml[p] * s1;
    }
    const float out = x.A[p] * u + x.B[p] * o + acc;
    float inj = 0.f;
    const int d = x.imap ? x.imap[p] : -1;
    if (d >= 0) {
        for (int c = x.ioff[d]; c < x.ioff[d + 1]; ++c) inj += x.icoef[c] * x.ivalues[x.ipoint[c]];
        inj *= x.isign;
    }
    x.out[p] = d >= 0 ? out + inj : out;
    if (x.gps && y >= g.nabs[1] && y < ny - g.nabs[1] && z >= g.nabs[2] && z < nz - g.nabs[2]) {
        // gradients on the interior, from the snapshot (box layout) and the adjoint level u
        const long long bi = sa_boxidx(x.box, 0, y, z);
        const float *ps = x.gps;
        if (x.ggv) {
            float pdt;
            if (x.graw) {
                const float pp = x.gpsp ? x.gpsp[bi] : 0.f, pm = x.gpsm ? x.gpsm[bi] : 0.f;
                pdt = (pp - 2.f * ps[bi] + pm) * x.gfdt2;
            } else {
                pdt = ps[bi];
            }
            x.ggv[p] += pdt * u;
            x.gpv[p] += pdt * pdt;
        }
        if (RHO == 2 && x.ggr) {
            const long long sy = x.box.p2;  // snapshot row pitch
            const float *bc = sB + (w + 5) * BZ + lane + 8;  // own cell in the b tile
            float acc2 = 0.f, lap = 0.f;
#pragma unroll
            for (int a = 1; a <= 2; ++a) {
                const long long st = a == 1 ? sy : 1;
                const int bs = a == 1 ? BZ : 1;
                float db = 0.f, dp = 0.f, l = g.w2[a][0] * ps[bi];
#pragma unroll
                for (int r = 0; r < 5; ++r) {
                    db += g.ws[a][r] * (bc[r * bs] - bc[-(r + 1) * bs]);
                    dp += g.ws[a][r] * (ps[bi + r * st] - ps[bi - (r + 1) * st]);
                }
#pragma unroll
                for (int r = 1; r <= 5; ++r) l += g.w2[a][r] * (ps[bi + r * st] + ps[bi - r * st]);
                acc2 += db * dp;
                lap += l;
            }
            const float f = -acc2 - bc[0] * lap;
            x.ggr[p] += f * u;
            x.gpr[p] += f * f;
        }
    }
    if (SAVE && sa_inbox(x.box, 0, y, z)) {
        float sv = x.save_mode == 1 ? (out - 2.f * u + o) * x.idt2 : u;
        if (d >= 0 && x.isave) sv += inj * x.idt2;
        x.save[sa_boxidx(x.box, 0, y, z)] = sv;
    }
