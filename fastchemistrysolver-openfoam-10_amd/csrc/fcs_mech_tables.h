// fcs_mech_tables.h -- host-side construction of the device tables of a
// mechanism (the work of fcs_mech_create, replacing OptReaction::readInfo,
// src/Reaction/OptReaction.C:348-1339): reactions ordered by rate form,
// per-species tables padded to SW, the Jacobian's term lists, the
// third-body entries and the general-path rows, all in one image whose
// placement the caller decides (device memory for the library, host
// memory for the host build of the device code used by the tests).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "fcs.h"
#include "fcs_device.h"

namespace fcs {

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Builds the tables of *d for the padded size NP (group width SW); place(host)
// must put the image somewhere the kernels can read and return its address
// (nullptr on failure).  dm receives the table pointers and sizes, perm the
// device order of the reactions (device index -> file index).
template <class Place>
int build_mech_tables(const fcs_mech_desc* d, int NP, int SW, Place&& place, DevMech& dm, std::vector<int>& perm)
{
    if (!d || d->n_species < 1 || d->n_reactions < 0) return FCS_E_ARG;
    const int ns = d->n_species, nr = d->n_reactions;
    // per-species tables are padded to the threads of one cell; ONE is the
    // concentration slot holding 1.0 (the "no other species" Jacobian factor)
    const int ONE = NP > 64 ? 255 : 64;
    // ---- order reactions by rate form (file order within a form)
    perm.clear();
    for (int t = 0; t <= FCS_RXN_SRI; ++t)
        for (int r = 0; r < nr; ++r)
            if (d->rxn_type[r] == t) perm.push_back(r);
    if ((int)perm.size() != nr) return FCS_E_ARG;
    // slot layout: up to 3 reactant slots, up to slot_stride product slots
    const int ss = d->slot_stride > 0 ? d->slot_stride : FCS_MAX_STOICH;
    if (ss < FCS_MAX_STOICH || ss > FCS_MAX_SLOTS) return FCS_E_ARG;
    for (int r = 0; r < nr; ++r) {
        const int nl = d->rxn_n_lhs[r], nrh = d->rxn_n_rhs[r];
        if (nl < 1 || nl > FCS_MAX_STOICH || nrh < 1 || nrh > ss) return FCS_E_ARG;
        for (int q = 0; q < nl; ++q)
            if (d->rxn_lhs[(size_t)ss * r + q] < 0 || d->rxn_lhs[(size_t)ss * r + q] >= ns) return FCS_E_ARG;
        for (int q = 0; q < nrh; ++q)
            if (d->rxn_rhs[(size_t)ss * r + q] < 0 || d->rxn_rhs[(size_t)ss * r + q] >= ns) return FCS_E_ARG;
    }
    const int pmax = d->plog_max_points > 0 ? d->plog_max_points : 1;
    // ---- host images of every table
    std::vector<double> W(SW, 0.0), invW(SW, 0.0), Tc(SW, 1000.0), clo(7 * SW, 0.0), chi(7 * SW, 0.0), Hf(SW, 0.0);
    double TlowMin = d->T_low[0], ThighMax = d->T_high[0], TcMin = d->T_common[0], TcMax = d->T_common[0];
    for (int s = 0; s < ns; ++s) {
        W[s] = d->W[s];
        invW[s] = 1.0 / d->W[s];
        Tc[s] = d->T_common[s];
        for (int k = 0; k < 7; ++k) {
            clo[k * SW + s] = d->coeffs_low[7 * s + k];
            chi[k * SW + s] = d->coeffs_high[7 * s + k];
        }
        TlowMin = std::min(TlowMin, d->T_low[s]);
        ThighMax = std::max(ThighMax, d->T_high[s]);
        TcMin = std::min(TcMin, d->T_common[s]);
        TcMax = std::max(TcMax, d->T_common[s]);
        // janafThermo::Hf of OpenFOAM-10 (not in the reference tree): the
        // low-range fit's enthalpy at Tstd = 298.15 K, per unit mass
        const double* a = d->coeffs_low + 7 * s;
        const double Ts = 298.15;
        Hf[s] = (RU / d->W[s]) * (((((a[4] / 5 * Ts + a[3] / 4) * Ts + a[2] / 3) * Ts + a[1] / 2) * Ts + a[0]) * Ts + a[5]);
    }
    std::vector<char> slotUnused(6 * (size_t)std::max(nr, 1), 1);
    std::vector<int> rtype(nr), rrev(nr), rnl(nr), rnr(nr), rsp(6 * std::max(nr, 1), 0), rtb(nr, -1), raux(nr, -1);
    std::vector<double> rA(nr), rbeta(nr), rTa(nr), rA0(nr), rbeta0(nr), rTa0(nr);
    std::vector<double> troe, eff, plogLnp, plogLnA, plogBeta, plogTa, rtroe(5 * (size_t)std::max(nr, 1), 0.0);
    std::vector<int> plogN;
    int nTB = 0;
    for (int r = 0; r < nr; ++r) {
        const int f = perm[r];
        const int t = d->rxn_type[f];
        rtype[r] = t;
        rrev[r] = d->rxn_reversible[f] == 2 ? 2 : (d->rxn_reversible[f] ? 1 : 0);
        rnl[r] = d->rxn_n_lhs[f];
        rnr[r] = d->rxn_n_rhs[f];
        // the reference evaluates reactions with more than 3 slots on a side
        // through RFG / JFG (ReactionRate.C:565-569, ReactionFunctionI.H:
        // 648-720, JacobianFunctionI.H:2965-3060), where a nonEquilibrium row
        // has no reverse rate; the products beyond slot 3 then enter only
        // through the net stoichiometry (nu below) and the slot count
        if (rnr[r] > FCS_MAX_STOICH && rrev[r] == 2) rrev[r] = 0;
        // a Kc reverse rate over such a side takes every product slot
        // (the device's general branch, rrev 3)
        if (rnr[r] > FCS_MAX_STOICH && rrev[r] == 1) rrev[r] = 3;
        for (int q = 0; q < 3; ++q) {
            rsp[q * nr + r] = q < rnl[r] ? d->rxn_lhs[(size_t)ss * f + q] : 0;
            rsp[(3 + q) * nr + r] = q < rnr[r] ? d->rxn_rhs[(size_t)ss * f + q] : 0;
            slotUnused[q * nr + r] = q >= rnl[r];
            slotUnused[(3 + q) * nr + r] = q >= rnr[r];
        }
        rA[r] = d->A[f];
        rbeta[r] = d->beta[f];
        rTa[r] = d->Ta[f];
        rA0[r] = d->A0 ? d->A0[f] : 0.0;
        rbeta0[r] = d->beta0 ? d->beta0[f] : 0.0;
        rTa0[r] = d->Ta0 ? d->Ta0[f] : 0.0;
        if (rrev[r] == 2) {
            // nonEquilibrium (Arrhenius and third-body forms only): the k_0
            // slot, unused by these forms, carries the reverse Arrhenius
            if ((t != FCS_RXN_ARRHENIUS && t != FCS_RXN_THIRD_BODY) || !d->A_rev || !d->beta_rev || !d->Ta_rev)
                return FCS_E_ARG;
            rA0[r] = d->A_rev[f];
            rbeta0[r] = d->beta_rev[f];
            rTa0[r] = d->Ta_rev[f];
        }
        if (t == FCS_RXN_THIRD_BODY || t == FCS_RXN_LINDEMANN || t == FCS_RXN_TROE || t == FCS_RXN_SRI) {
            if (!d->efficiencies) return FCS_E_ARG;
            rtb[r] = nTB++;
            for (int s = 0; s < SW; ++s) eff.push_back(s < ns ? d->efficiencies[(size_t)f * ns + s] : 0.0);
        }
        if (t == FCS_RXN_TROE) {
            if (!d->troe) return FCS_E_ARG;
            raux[r] = (int)troe.size() / 4;
            for (int q = 0; q < 4; ++q) troe.push_back(d->troe[4 * (size_t)f + q]);
            for (int q = 0; q < 4; ++q) rtroe[(size_t)q * nr + r] = d->troe[4 * (size_t)f + q];
        }
        if (t == FCS_RXN_SRI) {
            if (!d->sri) return FCS_E_ARG;
            for (int q = 0; q < 5; ++q) rtroe[(size_t)q * nr + r] = d->sri[5 * (size_t)f + q];
        }
        if (t == FCS_RXN_PLOG) {
            const int np = d->plog_n[f];
            if (np < 1 || np > pmax) return FCS_E_ARG;
            raux[r] = (int)plogN.size();
            plogN.push_back(np);
            for (int q = 0; q < pmax; ++q) {
                const int qq = q < np ? q : np - 1;
                const size_t k = (size_t)f * pmax + qq;
                plogLnp.push_back(std::log(d->plog_p[k]));
                plogLnA.push_back(std::log(d->plog_A[k]));
                plogBeta.push_back(d->plog_beta[k]);
                plogTa.push_back(d->plog_Ta[k]);
            }
        }
    }
    // PLOG rows: contiguous in device order (sorted by form), row = r - plogBase
    int plogBase = nr;
    for (int r = 0; r < nr; ++r)
        if (rtype[r] == FCS_RXN_PLOG) { plogBase = r; break; }
    for (int r = 0; r < nr; ++r)
        if (rtype[r] == FCS_RXN_PLOG && raux[r] != r - plogBase) return FCS_E_ARG;
    // the device fetches the k_0 / reverse Arrhenius triples from extLo on and
    // the Troe / SRI coefficients from troeLo on
    int extLo = nr, troeLo = nr;
    for (int r = nr - 1; r >= 0; --r) {
        const int t = rtype[r];
        if (t == FCS_RXN_LINDEMANN || t == FCS_RXN_TROE || t == FCS_RXN_SRI || rrev[r] == 2) extLo = r;
        if (t == FCS_RXN_TROE || t == FCS_RXN_SRI) troeLo = r;
    }
    // general-path rows of the rrev 3 reactions: slot count, LDS index of the
    // first reverse partial (the reaction arrays continue past nr), species
    std::vector<int> rgen(std::max(nr, 1), -1), gen;
    int nrq = nr;
    for (int r = 0; r < nr; ++r) {
        if (rrev[r] != 3) continue;
        rgen[r] = (int)gen.size() / GEN_STRIDE;
        gen.push_back(rnr[r]);
        gen.push_back(nrq);
        for (int q = 0; q < GEN_STRIDE - 2; ++q) gen.push_back(q < rnr[r] ? d->rxn_rhs[(size_t)ss * perm[r] + q] : 0);
        nrq += rnr[r];
    }
    if (nrq > 0xffff) return FCS_E_ARG;   // 16-bit reaction field of the Jacobian terms
    // third_body_sums reads rows t0 + [0, 2 SW) per pass unconditionally
    const int nTBpad = std::max(2 * SW, (nTB + 2 * SW - 1) / (2 * SW) * (2 * SW));
    std::vector<double> effT((size_t)SW * nTBpad, 0.0);
    for (int t = 0; t < nTB; ++t)
        for (int s = 0; s < SW; ++s) effT[(size_t)s * nTBpad + t] = eff[(size_t)t * SW + s];
    // ---- per-species lists (net stoichiometry), lane-interleaved
    std::vector<std::vector<std::pair<int, double>>> lst(ns), lstTB(ns);
    for (int r = 0; r < nr; ++r) {
        std::vector<double> nu(ns, 0.0);
        const int f = perm[r];
        for (int q = 0; q < rnl[r]; ++q) nu[rsp[q * nr + r]] -= 1.0;
        for (int q = 0; q < rnr[r]; ++q) nu[d->rxn_rhs[(size_t)ss * f + q]] += 1.0;
        for (int s = 0; s < ns; ++s) {
            if (nu[s] == 0.0) continue;
            lst[s].push_back({r, nu[s]});
            if (rtb[r] >= 0) lstTB[s].push_back({rtb[r], nu[s]});
        }
    }
    int maxDeg = 1;
    for (int s = 0; s < ns; ++s) maxDeg = std::max(maxDeg, (int)lst[s].size());
    maxDeg = (maxDeg + 15) / 16 * 16;   // the device walks the lists in unconditional blocks of 16
    // packed reaction species: slots 0-2 left, 3-5 right (bytes 0-5), nl, nr (bytes 6, 7)
    std::vector<unsigned long long> rpack(std::max(nr, 1), 0ull);
    for (int r = 0; r < nr; ++r) {
        unsigned long long w = 0;
        for (int q = 0; q < 6; ++q) {
            // the wavefront form names its identity slot (64) in unused slots
            const int sq = (NP <= 64 && slotUnused[q * nr + r]) ? 64 : rsp[q * nr + r];
            w |= (unsigned long long)(sq & 255) << (8 * q);
        }
        w |= (unsigned long long)rnl[r] << 48;
        w |= (unsigned long long)rnr[r] << 56;
        rpack[r] = w;
    }
    std::vector<unsigned long long> spPack((size_t)maxDeg * SW, 0ull);
    std::vector<int> spRN((size_t)maxDeg * SW, 8 << 16);   // padding: reaction 0, nu 0
    std::vector<int> spDeg(SW, 0);
    for (int s = 0; s < ns; ++s) spDeg[s] = (int)lst[s].size();
    std::vector<int> tbzRow, tbzTb;
    std::vector<double> tbzNu;
    for (int s = 0; s < ns; ++s) {
        for (size_t e = 0; e < lst[s].size(); ++e) {
            const int nuI = (int)lst[s][e].second;
            const int r = lst[s][e].first;
            if (nuI != lst[s][e].second || nuI < -7 || nuI > 7 || r > 0xffff) return FCS_E_ARG;
            spRN[e * SW + s] = r | ((nuI + 8) << 16);
            unsigned long long pk = rpack[r];
            if (rrev[r] == 3) {
                // the Jacobian round stores kr times the whole product side in
                // rk: the entry names no product slot (identity slots, count 0)
                pk &= ~(0xffffffull << 24 | 0xffull << 56);
                if (NP <= 64) pk |= 64ull << 24 | 64ull << 32 | 64ull << 40;
            }
            spPack[e * SW + s] = pk;
        }
        for (size_t e = 0; e < lstTB[s].size(); ++e) {
            tbzRow.push_back(s);
            tbzTb.push_back(lstTB[s][e].first);
            tbzNu.push_back(lstTB[s][e].second);
        }
    }
    // ---- one device allocation
    struct Piece { const void* src; size_t bytes; size_t off; };
    std::vector<Piece> pieces;
    size_t total = 0;
    auto add = [&](const void* src, size_t bytes) {
        pieces.push_back({src, bytes, total});
        total += align256(bytes > 0 ? bytes : 8);
        return pieces.size() - 1;
    };
    auto vd = [&](const std::vector<double>& v) { return add(v.data(), v.size() * sizeof(double)); };
    auto vi = [&](const std::vector<int>& v) { return add(v.data(), v.size() * sizeof(int)); };
    const size_t iW = vd(W), iInvW = vd(invW), iTc = vd(Tc), iClo = vd(clo), iChi = vd(chi), iHf = vd(Hf);
    // rate form, reverse kind and third-body row of a reaction in one word
    // (DevMech::rmeta)
    std::vector<int> rmeta(std::max(nr, 1), 0);
    for (int r = 0; r < nr; ++r) {
        if (rtb[r] + 1 >= (1 << 24)) return FCS_E_ARG;
        rmeta[r] = (int)((unsigned)(rtb[r] + 1) | ((unsigned)rtype[r] << 24) | ((unsigned)rrev[r] << 28));
    }
    const size_t iMeta = vi(rmeta), iNl = vi(rnl), iNr = vi(rnr), iSp = vi(rsp), iAux = vi(raux);
    const size_t iA = vd(rA), iB = vd(rbeta), iTa = vd(rTa), iA0 = vd(rA0), iB0 = vd(rbeta0), iTa0 = vd(rTa0);
    const size_t iTroe = vd(troe), iEff = vd(eff), iEffT = vd(effT), iRtroe = vd(rtroe);
    const size_t iPn = vi(plogN), iPp = vd(plogLnp), iPA = vd(plogLnA), iPb = vd(plogBeta), iPT = vd(plogTa);
    // per-species Jacobian term lists (nu_ir dq_r/dC_j per reactant / product
    // slot), sorted by column j (stable: reaction-list order within a column)
    std::vector<std::vector<unsigned long long>> terms(ns);
    for (int s = 0; s < ns; ++s) {
        for (const auto& e : lst[s]) {
            const int r = e.first, nu = (int)e.second;
            for (int side = 0; side < 2; ++side) {
                if (side == 1 && !rrev[r]) continue;
                if (side == 1 && rrev[r] == 3) {
                    // one term per product slot q: -nu times the partial
                    // kr prod_{b != q} C_b the Jacobian round leaves at LDS
                    // index gen[1] + q of the reaction arrays
                    const int* g = gen.data() + (size_t)rgen[r] * GEN_STRIDE;
                    for (int q = 0; q < g[0]; ++q)
                        terms[s].push_back((unsigned long long)(g[1] + q) | (unsigned long long)g[2 + q] << 16 |
                                           (unsigned long long)s << 24 | (unsigned long long)ONE << 32 |
                                           (unsigned long long)ONE << 40 | (unsigned long long)(-nu + 8) << 48);
                    continue;
                }
                const int cnt = side == 0 ? rnl[r] : rnr[r];
                const int sgn = side == 0 ? nu : -nu;
                for (int q = 0; q < cnt; ++q) {
                    int oth[2] = {ONE, ONE}, no = 0;
                    for (int b = 0; b < cnt; ++b)
                        if (b != q) oth[no++] = rsp[(3 * side + b) * nr + r];
                    const unsigned long long col = (unsigned long long)rsp[(3 * side + q) * nr + r];
                    terms[s].push_back((unsigned long long)r | col << 16 | (unsigned long long)s << 24 |
                                       (unsigned long long)oth[0] << 32 | (unsigned long long)oth[1] << 40 |
                                       (unsigned long long)(sgn + 8) << 48 | (unsigned long long)side << 56);
                }
            }
        }
        std::stable_sort(terms[s].begin(), terms[s].end(), [](unsigned long long a, unsigned long long b) {
            return ((a >> 16) & 255u) < ((b >> 16) & 255u);
        });
    }
    // the terms of one element stay in one thread; the elements, in row and
    // then column order, are cut into SW runs of at most maxTerms terms
    // (the smallest such bound: walked in steps of 16, so the bound is a
    // multiple of 16)
    std::vector<std::pair<size_t, size_t>> elems;   // [first, last) in flat
    std::vector<unsigned long long> flat;
    for (int s = 0; s < ns; ++s)
        for (size_t e = 0; e < terms[s].size(); ++e) {
            if (e == 0 || ((terms[s][e] >> 16) & 255u) != ((terms[s][e - 1] >> 16) & 255u))
                elems.push_back({flat.size(), flat.size()});
            flat.push_back(terms[s][e]);
            elems.back().second = flat.size();
        }
    int maxTerms = 16;
    for (const auto& e : elems) maxTerms = std::max(maxTerms, (int)(e.second - e.first + 15) / 16 * 16);
    std::vector<size_t> runStart;   // first element of each thread's run
    for (;; maxTerms += 16) {
        runStart.assign(1, 0);
        size_t len = 0;
        for (size_t e = 0; e < elems.size(); ++e) {
            const size_t n = elems[e].second - elems[e].first;
            if (len + n > (size_t)maxTerms) {
                runStart.push_back(e);
                len = 0;
            }
            len += n;
        }
        if ((int)runStart.size() <= SW) break;
    }
    runStart.push_back(elems.size());
    // padding after a thread's last element repeats that element's key with
    // coefficient 0 (the same sum is stored again); a thread without an
    // element stores into its dump slot
    const unsigned long long pad = (unsigned long long)ONE << 32 | (unsigned long long)ONE << 40 | 8ull << 48;
    std::vector<unsigned long long> spTerm((size_t)maxTerms * SW, pad | 1ull << 57);
    for (size_t t = 0; t + 1 < runStart.size(); ++t) {
        const size_t f0 = elems[runStart[t]].first, f1 = runStart[t + 1] > runStart[t]
                                                         ? elems[runStart[t + 1] - 1].second : f0;
        for (size_t e = 0; e < (size_t)maxTerms; ++e)
            spTerm[e * SW + t] = f0 + e < f1 ? flat[f0 + e]
                                             : (f1 > f0 ? (flat[f1 - 1] & 0xffff0000ull) | pad : spTerm[e * SW + t]);
    }
    // flat third-body entry list padded to a multiple of 16 with zero entries on the last row
    const int nTBz = (int)tbzRow.size();
    while (tbzRow.size() % 16) {
        tbzRow.push_back(nTBz ? tbzRow.back() : 0);
        tbzTb.push_back(0);
        tbzNu.push_back(0.0);
    }
    const size_t iSpDeg = vi(spDeg);
    const size_t iSpRN = vi(spRN), iSpTerm = add(spTerm.data(), spTerm.size() * 8);
    // first entry of each 16-row block (rows ascending; the padding repeats the last row)
    std::vector<int> tbzBlk(SW / 16 + 1, (int)tbzRow.size());
    for (int b = SW / 16 - 1; b >= 0; --b) {
        int e = tbzBlk[b + 1];
        while (e > 0 && tbzRow[e - 1] >= 16 * b) --e;
        tbzBlk[b] = nTBz > 0 ? e : 0;
    }
    const size_t iTbzRow = vi(tbzRow), iTbzTb = vi(tbzTb), iTbzNu = vd(tbzNu), iTbzBlk = vi(tbzBlk);
    auto vq = [&](const std::vector<unsigned long long>& v) { return add(v.data(), v.size() * 8); };
    const size_t iRpack = vq(rpack), iSpPack = vq(spPack);
    const size_t iRgen = vi(rgen), iGen = vi(gen);
    std::vector<unsigned char> host(total, 0);
    for (const auto& p : pieces)
        if (p.bytes) std::memcpy(host.data() + p.off, p.src, p.bytes);
    char* base = (char*)place(host);
    if (!base) return FCS_E_NOMEM;
    auto P = [&](size_t i) { return (void*)(base + pieces[i].off); };
    dm.ns = ns;
    dm.nr = nr;
    dm.nrq = nrq;
    dm.extLo = extLo;
    dm.troeLo = troeLo;
    dm.plogBase = plogBase;
    dm.nPlog = (int)plogN.size();
    dm.rgen = (const int*)P(iRgen);
    dm.gen = (const int*)P(iGen);
    dm.NPm = SW;
    dm.nTB = nTB;
    dm.pmax = pmax;
    dm.maxDeg = maxDeg;
    dm.nTBpad = nTBpad;
    dm.TlowMin = TlowMin;
    dm.ThighMax = ThighMax;
    dm.TcMin = TcMin;
    dm.TcMax = TcMax;
    dm.W = (const double*)P(iW);
    dm.invW = (const double*)P(iInvW);
    dm.Tc = (const double*)P(iTc);
    dm.clo = (const double*)P(iClo);
    dm.chi = (const double*)P(iChi);
    dm.Hf = (const double*)P(iHf);
    dm.rnl = (const int*)P(iNl);
    dm.rnr = (const int*)P(iNr);
    dm.rsp = (const int*)P(iSp);
    dm.rmeta = (const int*)P(iMeta);
    dm.raux = (const int*)P(iAux);
    dm.rA = (const double*)P(iA);
    dm.rbeta = (const double*)P(iB);
    dm.rTa = (const double*)P(iTa);
    dm.rA0 = (const double*)P(iA0);
    dm.rbeta0 = (const double*)P(iB0);
    dm.rTa0 = (const double*)P(iTa0);
    dm.troe = (const double*)P(iTroe);
    dm.rtroe = (const double*)P(iRtroe);
    dm.eff = (const double*)P(iEff);
    dm.effT = (const double*)P(iEffT);
    dm.plog_n = (const int*)P(iPn);
    dm.plog_lnp = (const double*)P(iPp);
    dm.plog_lnA = (const double*)P(iPA);
    dm.plog_beta = (const double*)P(iPb);
    dm.plog_Ta = (const double*)P(iPT);
    dm.sp_rn = (const int*)P(iSpRN);
    dm.sp_deg = (const int*)P(iSpDeg);
    dm.maxTerms = maxTerms;
    dm.sp_term = (const unsigned long long*)P(iSpTerm);
    dm.nTBz = nTBz > 0 ? (int)tbzRow.size() : 0;
    dm.tbz_row = (const int*)P(iTbzRow);
    dm.tbz_blk = (const int*)P(iTbzBlk);
    dm.tbz_tb = (const int*)P(iTbzTb);
    dm.tbz_nu = (const double*)P(iTbzNu);
    dm.rpack = (const unsigned long long*)P(iRpack);
    dm.sp_pack = (const unsigned long long*)P(iSpPack);
    return FCS_OK;
}

}  // namespace fcs
