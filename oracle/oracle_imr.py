"""ORACLE — test infrastructure only.

CPU restatement (numpy, plus the oracle's restatement of the reference's
spline in ``oracle_eob.cpp``) of the merger-ringdown attachment of
SEOBNRv4PHM.compute_full_waveform, eob/seobnrv4phm.py:2275-2560:

* NR peak fits (eob/seobnrv4phm.py:878-1462), the NQC solve
  (EOBCalculateNQCCoefficientsV4_freeattach, :1465-1683) and the NQC factor
  (EOBNonQCCorrection, :1686-1700);
* the ringdown fits and constrained coefficients (eob/eobstuff.py:44-342) and
  the ringdown ansatz (compute_MR_mode_free, eob/seobnrv4phm.py:1779-1844);
* the remnant mass (eob/eobstuff.py:6-19, SEOBNRv4's fit on the Kerr ISCO);
* the remnant spin and the QNM frequencies, which the reference takes from
  pesummary (HBR2016, "M3J4") and lalsimulation (tabulated QNMs).  Neither is in
  this image or in the reference tree: the spin is the published HBR2016
  formula and the QNMs come from Leaver's continued fractions (solved here
  with a Newton iteration over the whole spin path, an independent route from
  the device's secant refinement).  Parity with pesummary / lalsimulation is
  unpinned; the pins are the closed forms the tests check (ISCO at a = 0 and
  a -> 1, Schwarzschild QNMs, the small-spin angular eigenvalue, the
  equal-mass non-spinning remnant spin 0.6865).

Only ``tests/`` import this module.
"""
from __future__ import annotations

import math

import numpy as np

from . import oracle as _orc

MTSUN_SI = 4.925491025543576e-06
RD_MODES = ((2, 2), (2, 1), (3, 3), (4, 4), (5, 5))


# ---------------------------------------------------------------- NR fits
def _eta(m1, m2):
    return m1 * m2 / ((m1 + m2) * (m1 + m2))


def nr_delta_t22(m1, m2, chi1, chi2):
    """EOBGetNRSpinPeakDeltaTv4 for (2, 2), eob/seobnrv4phm.py:878-929."""
    eta = m1 * m2 / (m1 + m2) / (m1 + m2)
    chi = 0.5 * (chi1 + chi2) + 0.5 * (chi1 - chi2) * (m1 - m2) / (m1 + m2) / (1.0 - 2.0 * eta)
    terms = {(0, 0): 2.50499, (0, 1): 13.0064, (0, 2): 11.5435, (1, 0): 45.8838, (1, 1): -40.3183, (1, 3): -19.0538,
             (2, 0): 13.0879, (2, 3): 0.192775, (3, 0): -716.044}
    return sum(c * eta ** i * chi ** j for (i, j), c in terms.items())


def _combine(eta, A1, fEQ, fTPL):
    # eob/seobnrv4phm.py:932-942
    A0 = -0.00099601593625498 * A1 - 0.00001600025600409607 * fEQ + 1.000016000256004 * fTPL
    A2 = -3.984063745019967 * A1 + 16.00025600409607 * fEQ - 16.0002560041612 * fTPL
    return A0 + A1 * eta + A2 * eta * eta


def _swap(m1, m2, chi1, chi2):
    eta = _eta(m1, m2)
    dM = math.sqrt(1.0 - 4.0 * eta)
    chiS, chiA = 0.5 * (chi1 + chi2), 0.5 * (chi1 - chi2)
    if m1 < m2:
        m1, m2, chiA = m2, m1, -chiA
    chi = chiS + chiA * (m1 - m2) / (m1 + m2) / (1.0 - 2.0 * eta)
    return eta, dM, chiS, chiA, chi


def nr_amp(l, m, m1, m2, chi1, chi2):
    """EOBGetNRSpinPeakAmplitudeV4, eob/seobnrv4phm.py:945-1075."""
    eta, dM, chiS, chiA, chi = _swap(m1, m2, chi1, chi2)
    e2 = eta * eta
    if (l, m) == (2, 2):
        fTPL = 1.4528573105413543 + 0.16613449160880395 * chi + 0.027355646661735258 * chi ** 2 - 0.020072844926136438 * chi ** 3
        fEQ = 1.577457498227 - 0.0076949474494639085 * chi + 0.02188705616693344 * chi ** 2 + 0.023268366492696667 * chi ** 3
        A1 = -0.03442402416125921 - 1.218066264419839 * chi - 0.5683726304811634 * chi ** 2 + 0.4011143761465342 * chi ** 3
        return eta * _combine(eta, A1, fEQ, fTPL)
    if (l, m) == (2, 1):
        c = chiS * dM / (1.0 - 1.3 * eta) + chiA
        return -((0.29256703361640224 - 0.19710255145276584 * eta) * eta * c
                 + dM * eta * (-0.42817941710649793 + 0.11378918021042442 * eta - 0.7736772957051212 * e2
                               + c * c * (0.047004057952214004 - eta * 0.09326128322462478))
                 + dM * eta * c * (-0.010195081244587765 + 0.016876911550777807 * c * c))
    if (l, m) == (3, 3):
        c = chiS * dM + chiA
        return ((0.10109183988848384 * eta - 0.4704095462146807 * e2 + 1.0735457779890183 * e2 * eta) * c
                + dM * (0.5636580081367962 * eta - 0.054609013952480856 * e2 + 2.3093699480319234 * e2 * eta
                        + c * c * (0.029812986680919126 * eta - 0.09688097244145283 * e2)))
    if (l, m) == (4, 4):
        c = chiS * (1 - 5 * eta) + chiA * dM
        return (eta * (0.2646580063832686 + 0.067584186955327 * c + 0.02925102905737779 * c * c)
                + e2 * (-0.5658246076387973 - 0.8667455348964268 * c + 0.005234192027729502 * c * c)
                + eta * e2 * (-2.5008294352355405 + 6.880772754797872 * c - 1.0234651570264885 * c * c)
                + e2 * e2 * (7.6974501716202735 - 16.551524307203252 * c))
    if (l, m) == (5, 5):
        c = chiS * dM + chiA
        return (0.128621 * dM * eta - 0.474201 * dM * eta * eta + 1.0833 * dM * eta ** 3 + 0.0322784 * eta * c
                - 0.134511 * c * eta * eta + 0.0990202 * c * eta ** 3)
    raise ValueError((l, m))


def nr_adot(l, m, m1, m2, chi1, chi2):
    """EOBGetNRSpinPeakADotV4, eob/seobnrv4phm.py:1078-1158."""
    eta, dM, chiS, chiA, _ = _swap(m1, m2, chi1, chi2)
    e2 = eta * eta
    if (l, m) == (2, 2):
        return 0.0
    if (l, m) == (2, 1):
        c = chiS * dM / (1.0 - 2.0 * eta) + chiA
        return (dM * eta * (0.007147528020812309 - eta * 0.035644027582499495)
                + dM * eta * c * (-0.0087785131749995 + eta * 0.03054672006241107)
                + eta * 0.00801714459112299 * abs(-dM * (0.7875612917853588 + eta * 1.161274164728927 + e2 * 11.306060006923605) + c))
    if (l, m) == (3, 3):
        c = chiS * dM + chiA
        return (dM * eta * (-0.00309943555972098 + eta * 0.010076527264663805) * c * c
                + eta * 0.0016309606446766923 * math.sqrt(dM * dM * (8.811660714437027 + 104.47752236009688 * eta)
                                                          + dM * c * (-5.352043503655119 + eta * 49.68621807460999) + c * c))
    if (l, m) == (4, 4):
        c = chiS * (1 - 7 * eta) + chiA * dM
        return (eta * (0.004347588211099233 - 0.0014612210699052148 * c - 0.002428047910361957 * c * c)
                + e2 * (0.023320670701084355 - 0.02240684127113227 * c + 0.011427087840231389 * c * c)
                + eta * e2 * (-0.46054477257132803 + 0.433526632115367 * c) + e2 * e2 * (1.2796262150829425 - 1.2400051122897835 * c))
    if (l, m) == (5, 5):
        c = chiS * dM + chiA
        return eta * (dM * (-0.008389798844109389 + 0.04678354680410954 * eta) + dM * c * (-0.0013605616383929452 + 0.004302712487297126 * eta)
                      + dM * c * c * (-0.0011412109287400596 + 0.0018590391891716925 * eta)
                      + 0.0002944221308683548 * abs(dM * (37.11125499129578 - 157.79906814398277 * eta) + c))
    raise ValueError((l, m))


def nr_addot(l, m, m1, m2, chi1, chi2):
    """EOBGetNRSpinPeakADDotV4, eob/seobnrv4phm.py:1162-1258."""
    eta, dM, chiS, chiA, chi = _swap(m1, m2, chi1, chi2)
    e2 = eta * eta
    if (l, m) == (2, 2):
        x = chi - 1.0
        fTPL = 0.002395610769995033 * x - 0.00019273850675004356 * x * x - 0.00029666193167435337 * x ** 3
        fEQ = -0.004126509071377509 + 0.002223999138735809 * chi
        A1 = -0.005776537350356959 + 0.001030857482885267 * chi
        return eta * _combine(eta, A1, fEQ, fTPL)
    if (l, m) == (2, 1):
        c = chiS * dM / (1.0 - 2.0 * eta) + chiA
        return eta * dM * 0.00037132201959950333 - abs(
            dM * eta * (-0.0003650874948532221 - eta * 0.003054168419880019)
            + dM * eta * c * c * (-0.0006306232037821514 - eta * 0.000868047918883389 + e2 * 0.022306229435339213)
            + eta * c ** 3 * 0.0003402427901204342 + dM * eta * c * 0.00028398490492743)
    if (l, m) == (3, 3):
        c = chiS * dM + chiA
        return (dM * eta * (0.0009605689249339088 - 0.00019080678283595965 * eta) * c
                - 0.00015623760412359145 * eta * abs(dM * (4.676662024170895 + 79.20189790272218 * eta - 1097.405480250759 * e2
                                                           + 6512.959044311574 * eta * e2 - 13263.36920919937 * e2 * e2) + c))
    if (l, m) == (4, 4):
        return (eta * (-0.000301722928925693 + 0.0003215952388023551 * chi) + e2 * (0.006283048344165004 + 0.0011598784110553046 * chi)
                + e2 * eta * (-0.08143521096050622 - 0.013819464720298994 * chi) + e2 * e2 * (0.22684871200570564 + 0.03275749240408555 * chi))
    if (l, m) == (5, 5):
        c = chiS * dM + chiA
        return eta * (dM * (0.00012727220842255978 + 0.0003211670856771251 * eta) + dM * c * (-0.00006621677859895541 + 0.000328855327605536 * eta)
                      + c * c * (-0.00005824622885648688 + 0.00013944293760663706 * eta))
    raise ValueError((l, m))


def fit_chi(m1, m2, chi1, chi2):
    eta = _eta(m1, m2)
    return 0.5 * (chi1 + chi2) + 0.5 * (chi1 - chi2) * (m1 - m2) / (m1 + m2) / (1.0 - 2.0 * eta)


def nr_omega(l, m, eta, chi):
    """EOBGetNRSpinPeakOmegaV4, eob/seobnrv4phm.py:1261-1359."""
    e2 = eta * eta
    if (l, m) == (2, 2):
        A4 = 7.629921628648589 + 4 * (7.629921628648589 - 25.81979479453255) * (eta - 0.25)
        A3 = 10.26207326082448 + 4 * (10.26207326082448 - 25.85037751197443) * (eta - 0.25)
        return 0.5626787200433265 + (-0.08706198756945482 + 0.00174345193125868 * chi) * math.log(A3 - A4 * chi)
    if (l, m) == (2, 1):
        return (0.1743194440996283 + eta * 0.1938944514123048 + 0.1670063050527942 * e2 + 0.053508705425291826 * chi
                - eta * chi * 0.18460213023455802 + e2 * chi * 0.2187305149636044 + chi * chi * 0.030228846150378793
                - eta * chi * chi * 0.11222178038468673)
    if (l, m) == (3, 3):
        return (0.3973947703114506 + 0.16419332207671075 * chi + 0.1635531186118689 * chi ** 2 + 0.06140164491786984 * chi ** 3
                + eta * (0.6995063984915486 - 0.3626744855912085 * chi - 0.9775469868881651 * chi ** 2)
                + e2 * (-0.3455328417046369 + 0.31952307610699876 * chi + 1.9334166149686984 * chi ** 2))
    if (l, m) == (4, 4):
        return (0.5389359134370971 + 0.16635177426821202 * chi + 0.2075386047689103 * chi ** 2 + 0.15268115749910835 * chi ** 3
                + eta * (0.7617423831337586 + 0.009587856087825369 * chi - 1.302303785053009 * chi ** 2 - 0.5562751887042064 * chi ** 3)
                + e2 * (0.9675153069365782 - 0.22059322127958586 * chi + 2.678097398558074 * chi ** 2) - e2 * eta * 4.895381222514275)
    if (l, m) == (5, 5):
        return (0.6437545281817488 + 0.22315530037902315 * chi + 0.2956893357624277 * chi ** 2 + 0.17327819169083758 * chi ** 3
                + eta * (-0.47017798518175785 - 0.3929010618358481 * chi - 2.2653368626130654 * chi ** 2 - 0.5512998466154311 * chi ** 3)
                + e2 * (2.311483807604238 + 0.8829339243493562 * chi + 5.817595866020152 * chi ** 2))
    raise ValueError((l, m))


def nr_omegadot(l, m, eta, chi):
    """EOBGetNRSpinPeakOmegaDotV4, eob/seobnrv4phm.py:1363-1462."""
    e2 = eta * eta
    if (l, m) == (2, 2):
        fTPL = -0.011209791668428353 + (0.0040867958978563915 + 0.0006333925136134493 * chi) * math.log(68.47466578100956 - 58.301487557007206 * chi)
        fEQ = 0.01128156666995859 + 0.0002869276768158971 * chi
        A1 = 0.01574321112717377 + 0.02244178140869133 * chi
        return _combine(eta, A1, fEQ, fTPL)
    if (l, m) == (2, 1):
        return (0.0070987396362959514 + eta * 0.024816844694685373 - e2 * 0.050428973182277494 + eta * e2 * 0.03442040062259341
                - chi * 0.0017751850002442097 + eta * chi * 0.004244058872768811 - e2 * chi * 0.031996494883796855
                - chi * chi * 0.0035627260615894584 + eta * chi * chi * 0.01471807973618255 - chi ** 3 * 0.0019020967877681962)
    if (l, m) == (3, 3):
        return (0.010337157192240338 - 0.0053067782526697764 * chi ** 2 - 0.005087932726777773 * chi ** 3
                + eta * (0.027735564986787684 + 0.018864151181629343 * chi + 0.021754491131531044 * chi ** 2 + 0.01785477515931398 * chi ** 3)
                + e2 * (0.018084233854540898 - 0.08204268775495138 * chi))
    if (l, m) == (4, 4):
        return (0.013997911323773867 - 0.0051178205260273574 * chi - 0.0073874256262988 * chi ** 2
                + eta * (0.0528489379269367 + 0.01632304766334543 * chi + 0.02539072293029433 * chi ** 2)
                + e2 * (-0.06529992724396189 + 0.05782894076431308 * chi))
    if (l, m) == (5, 5):
        return (0.01763430670755021 - 0.00024925743340389135 * chi - 0.009240404217656968 * chi ** 2 - 0.007907831334704586 * chi ** 3
                + eta * (-0.1366002854361568 + 0.0561378177186783 * chi + 0.16406275673019852 * chi ** 2 + 0.07736232247880881 * chi ** 3)
                + e2 * (0.9875890632901151 - 0.31392112794887855 * chi - 0.5926145463423832 * chi ** 2) - 1.6943356548192614 * e2 * eta)
    raise ValueError((l, m))


# ------------------------------------------------------------ ringdown fits
# eob/eobstuff.py:44-322: per mode {(i, j): k} for the term k eta^i chi^j
_RD = {
    "c1f": {(2, 2): {(0, 0): 0.0830664, (0, 1): -0.0196758, (0, 2): -0.0136459, (1, 0): 0.0612892, (1, 1): 0.00146142, (2, 0): -0.0893454},
            (2, 1): {(0, 0): 0.07780330893915006, (0, 1): -0.05070638166864379, (1, 0): 0.24091006920322164, (1, 1): 0.38582622780596576,
                     (2, 0): -0.7456327888190485, (2, 1): -0.9695534075470388},
            (3, 3): {(0, 0): 0.07638733045623343, (0, 1): -0.030993441267953236, (1, 0): 0.2543447497371546, (1, 1): 0.2516879591102584,
                     (2, 0): -1.0892686061231245, (2, 1): -0.7980907313033606},
            (4, 4): {(0, 0): -0.06392710223439678, (0, 1): -0.03646167590514318, (1, 0): 0.345195277237925, (1, 1): 1.2777441574118558,
                     (2, 0): -1.764352185878576, (2, 1): -14.825262897834696, (3, 1): 40.67135475479875},
            (5, 5): {(0, 0): -0.06704614393611373, (0, 1): 0.021905949257025503, (1, 0): -0.24754936787743445, (1, 1): -0.0943771022698497,
                     (2, 0): 0.7588042862093705, (2, 1): 0.4357768883690394}},
    "c2f": {(2, 2): {(0, 0): -0.623953, (0, 1): -0.371365, (1, 0): 1.39777, (1, 1): 2.40203, (2, 0): -1.82173, (2, 1): -5.25339},
            (2, 1): {(0, 0): -1.2451852641667298, (0, 1): -1.195786238319961, (1, 0): 6.134202504312409, (1, 1): 15.66696619631313,
                     (2, 0): -14.67251127485556, (2, 1): -44.41982201432511},
            (3, 3): {(0, 0): -0.8325292359346013, (0, 1): -0.598880303198448, (1, 0): 2.767989795032018, (1, 1): 5.904371617277156,
                     (2, 0): -7.028151926115957, (2, 1): -18.232606706124482},
            (4, 4): {(0, 0): 0.7813275473485185, (0, 1): 0.8094706044462984, (1, 0): -5.18689829943586, (1, 1): -5.38343327318501,
                     (2, 0): 14.026415859369477, (2, 1): 0.1051625997942299, (3, 1): 46.978434956814006},
            (5, 5): {(0, 0): 1.6763424265367357, (0, 1): 0.4925695499534606, (1, 0): -5.604559311983177, (1, 1): -6.209095657439377,
                     (2, 0): 16.751319143123386, (2, 1): 16.778452555342554}},
    "d1f": {(2, 2): {(0, 0): 0.147584, (0, 1): 0.00779176, (0, 2): -0.0244358, (1, 0): 0.263456, (1, 1): -0.120853, (2, 0): -0.808987},
            (2, 1): {(0, 0): 0.15601401627613815, (0, 1): 0.10219957917717622, (1, 0): 0.023346852452720928, (1, 1): -0.9435308286367039,
                     (2, 0): 0.15326558178697175, (2, 1): 1.7979082057513565},
            (3, 3): {(0, 0): 0.11085299117493969, (0, 1): 0.018959099261813613, (1, 0): 0.9999800463662053, (1, 1): -0.729149797691242,
                     (2, 0): -3.3983315694441125, (2, 1): 2.5192011762934037},
            (4, 4): {(0, 0): 0.11498976343440313, (0, 1): 0.008389519706605305, (1, 0): 1.6126522800609633, (1, 1): -0.8069979888526699,
                     (2, 0): -6.255895564079467, (2, 1): 7.595651881827078, (3, 1): -19.32367406125053},
            (5, 5): {(0, 0): 0.16465380962882128, (0, 1): -0.026574817803812007, (1, 0): -0.19184523794508765, (1, 1): -0.05519618962738479,
                     (2, 0): 0.33328424135336965, (2, 1): 0.3194274548351241}},
    "d2f": {(2, 2): {(0, 0): 2.46654, (0, 1): 3.13067, (0, 2): 0.581626, (1, 0): -6.99396, (1, 1): -9.61861, (2, 0): 17.5646},
            (2, 1): {(0, 0): 2.7886287922318105, (0, 1): 4.29290053494256, (1, 0): -0.8145406685320334, (1, 1): -15.93796979597706,
                     (2, 0): 5.549338798935832, (2, 1): 12.649775582333442, (0, 2): 2.5582321247274726, (1, 2): -10.232928498772893},
            (3, 3): {(0, 0): 2.7825237371542735, (0, 1): 2.8796835808075003, (1, 0): -7.844741660437831, (1, 1): -34.7670039322078,
                     (2, 0): 27.181024362399302, (2, 1): 127.13948436435182},
            (4, 4): {(0, 0): 3.111817347262856, (0, 1): 5.399341180960216, (1, 0): 15.885333959709488, (1, 1): -87.92421137153823,
                     (2, 0): -79.64931908155609, (2, 1): 657.7156442271963, (3, 1): -1555.2968529739226, (0, 2): 2.3832321567874686,
                     (1, 2): -9.532928476043567},
            (5, 5): {(0, 0): 11.102447263357977, (0, 1): 6.015112119742853, (1, 0): -58.605776859097084, (1, 1): -81.68032025902797,
                     (2, 0): 176.60643662729498, (2, 1): 266.47345742836745}},
}


def rd_fit(name, l, m, eta, chi):
    """One of c1f, c2f, d1f, d2f (EOBCalculateRD{Amplitude,Phase}Coefficient{1,2})."""
    return sum(k * eta ** i * chi ** j for (i, j), k in _RD[name][(l, m)].items())


# ------------------------------------------------------------------ remnant
def kerr_isco_radius(a):
    z1 = 1.0 + np.cbrt(1.0 - a * a) * (np.cbrt(1.0 + a) + np.cbrt(1.0 - a))
    z2 = np.sqrt(3.0 * a * a + z1 * z1)
    return 3.0 + z2 - np.sign(a) * np.sqrt((3.0 - z1) * (3.0 + z1 + 2.0 * z2))


def final_mass(m1, m2, a1, a2):
    """compute_final_mass_SEOBNRv4, eob/eobstuff.py:6-19 (lalsimulation's Kerr
    ISCO radius and E_isco = sqrt(1 - 2/(3 r)) written out)."""
    q = m1 / m2
    eta = m1 * m2 / (m1 + m2) ** 2
    iq = 1.0 / q
    atl = (a1 + a2 * iq * iq) / ((1.0 + iq) ** 2)
    r = kerr_isco_radius(atl)
    e = np.sqrt(1.0 - 2.0 / (3.0 * r))
    return 1.0 - ((1.0 - e) * eta + 16.0 * eta * eta * (0.00258 - 0.0773 / (1.0 / ((1.0 + iq * iq) / (1.0 + iq) ** 2) * atl - 1.6939)
                                                       - 0.25 * (1.0 - e)))


_HBR_K = np.array([[-5.97723, 3.39221, 4.48865, -5.77101, -13.0459],
                   [35.1278, -72.9336, -86.0036, 93.7371, 200.975],
                   [-146.822, 387.184, 447.009, -467.383, -884.339],
                   [223.911, -648.502, -697.177, 753.738, 1166.89]])
_HBR_XI = 0.474046


def final_spin(m1, m2, chi1, chi2):
    """HBR2016 remnant spin, n_M = 3, n_J = 4 (Hofmann, Barausse & Rezzolla 2016)."""
    M = m1 + m2
    nu = m1 * m2 / M ** 2
    atot = (chi1 * m1 ** 2 + chi2 * m2 ** 2) / M ** 2
    aeff = atot + _HBR_XI * nu * (chi1 + chi2)
    r = kerr_isco_radius(aeff)
    L = 2.0 / (3.0 * np.sqrt(3.0)) * (1.0 + 2.0 * np.sqrt(3.0 * r - 2.0))
    E = np.sqrt(1.0 - 2.0 / (3.0 * r))
    s = sum(_HBR_K[i, j] * nu ** (2 + i) * aeff ** j for i in range(4) for j in range(5))
    return atot + nu * (L - 2.0 * atot * (E - 1.0)) + s


# ---------------------------------------------------------------- QNMs
def _angular_cf(A, c, s, m, depth=80):
    k1, k2 = abs(m - s) / 2.0, abs(m + s) / 2.0
    t = None
    for n in range(depth, -1, -1):
        beta = (n * (n - 1) + 2 * n * (k1 + k2 + 1 - 2 * c) - (2 * c * (2 * k1 + s + 1) - (k1 + k2) * (k1 + k2 + 1))
                - (c * c + s * (s + 1) + A))
        if t is None:
            t = beta
            continue
        alpha = -2.0 * (n + 1) * (n + 2 * k1 + 1)
        gamma = 2 * c * (n + 1 + k1 + k2 + s)
        t = beta - alpha * gamma / t
    return t


def _newton(f, x, tol=1e-14, it=80, h=1e-7):
    for _ in range(it):
        fx = f(x)
        d = (f(x + h) - f(x - h)) / (2 * h)
        dx = fx / d
        x = x - dx
        if abs(dx) <= tol * (1 + abs(x)):
            break
    return x


def _radial_cf(w, a, s, m, A, depth=1200):
    b = math.sqrt(1.0 - 4.0 * a * a)
    wm = w / 2 - a * m
    c0 = 1 - s - 1j * w - 2j / b * wm
    c1 = -4 + 2j * w * (2 + b) + 4j / b * wm
    c2 = s + 3 - 3j * w - 2j / b * wm
    c3 = w * w * (4 + 2 * b - a * a) - 2 * a * m * w - s - 1 + (2 + b) * 1j * w - A + (4 * w + 2j) / b * wm
    c4 = s + 1 - 2 * w * w - (2 * s + 3) * 1j * w - (4 * w + 2j) / b * wm
    t = -2.0 * depth * depth + (c1 + 2) * depth + c3
    for n in range(depth - 1, -1, -1):
        alpha = n * n + (c0 + 1) * n + c0
        gamma = (n + 1) ** 2 + (c2 - 3) * (n + 1) + c4 - c2 + 2
        beta = -2.0 * n * n + (c1 + 2) * n + c3
        t = beta - alpha * gamma / t
    return t


def kerr_qnm(l, m, af, step=0.01):
    """M omega_lm0 of a Kerr hole of spin af (M = 1, e^{-i omega t}, Im < 0),
    Leaver's method followed from a = 0 in steps of ``step``."""
    s = -2
    schw = {2: 0.3736716844 - 0.0889623157j, 3: 0.5994432884 - 0.0927030479j, 4: 0.8091783848 - 0.0941639638j,
            5: 1.0122989967 - 0.0948727131j}
    w = 2 * schw[l]
    A = complex((l - s) * (l + s + 1))
    n = 1 + int(abs(af) / step)
    for k in range(1, n + 1):
        a = 0.5 * af * k / n
        state = {"A": A}

        def f(x):
            state["A"] = _newton(lambda y: _angular_cf(y, a * x, s, m), state["A"])
            return _radial_cf(x, a, s, m, state["A"])

        w = _newton(f, w)
        A = state["A"]
    return w / 2


def qnm_lal(l, m, af, mf):
    """The QNM frequency the reference uses (lalsimulation's convention:
    omega_r + i / tau), in units of the total mass."""
    w = kerr_qnm(l, m, af)
    return np.conj(w) / mf


# ------------------------------------------------------------ the attachment
def spline1(t, y):
    """The reference's not-a-knot spline of one series (oracle restatement of
    src/interpolate.cu): coefficients (y, c1, c2, c3)."""
    t = np.ascontiguousarray(t, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    n = t.size
    B, c1, c2, c3 = _orc.interpolate_TD(t.reshape(n, 1), y.reshape(n, 1), np.array([n], np.int32), 1, 1)
    return t, y, c1.ravel(), c2.ravel(), c3.ravel()


def seval(sp, T, deriv=0):
    t, y, c1, c2, c3 = sp
    return _orc.spline_eval(np.atleast_1d(np.asarray(T, dtype=np.float64)), t, y, c1, c2, c3, t.size, deriv)


def _argmin_first(a):
    return int(np.argmin(a))


def nqc_coefficients(l, m, t, amp, phase, r, pr, om, npts, t_peak, dT_nqc, m1, m2, chi1, chi2):
    """One (binary, mode) of EOBCalculateNQCCoefficientsV4_freeattach: t, amp,
    phase, r, pr, om on the uniform grid (npts valid samples), dT_nqc the mode's
    nrDeltaT.  Returns (a1, a2, a3, b1, b2)."""
    T = t_peak - dT_nqc
    idx = _argmin_first(np.abs(t - T))
    N = 7
    lo, hi = idx - N, npts
    tf = t[lo:hi]
    rO = r[lo:hi] * om[lo:hi]
    q1 = pr[lo:hi] ** 2 / rO ** 2
    q2 = q1 / r[lo:hi]
    q3 = q2 / np.sqrt(r[lo:hi])
    p1 = pr[lo:hi] / rO
    p2 = p1 * pr[lo:hi] ** 2
    a = amp[lo:hi]
    S = {k: spline1(tf, v) for k, v in (("q1", q1 * a), ("q2", q2 * a), ("q3", q3 * a), ("amp", a), ("p1", p1), ("p2", p2),
                                         ("ph", phase[lo:hi]))}
    Q = np.array([[seval(S[k], T, i)[0] for k in ("q1", "q2", "q3")] for i in range(3)])
    eta = _eta(m1, m2)
    rhs = np.array([abs(nr_amp(l, m, m1, m2, chi1, chi2)) - seval(S["amp"], T, 0)[0],
                    nr_adot(l, m, m1, m2, chi1, chi2) - seval(S["amp"], T, 1)[0],
                    nr_addot(l, m, m1, m2, chi1, chi2) - seval(S["amp"], T, 2)[0]])
    a123 = np.linalg.solve(Q, rhs)
    P = np.array([[-seval(S["p1"], T, i)[0], -seval(S["p2"], T, i)[0]] for i in (1, 2)])
    w, wd = seval(S["ph"], T, 1)[0], seval(S["ph"], T, 2)[0]
    omega = abs(w)
    omegaDot = abs(wd) if w * wd > 0.0 else -abs(wd)
    chi = fit_chi(m1, m2, chi1, chi2)
    b12 = np.linalg.solve(P, np.array([nr_omega(l, m, eta, chi) - omega, nr_omegadot(l, m, eta, chi) - omegaDot]))
    return np.array([a123[0], a123[1], a123[2], b12[0], b12[1]])


def nqc_factor(c, r, pr, om):
    """EOBNonQCCorrection with a3S = a4 = a5 = b3 = b4 = 0 (:1686-1700)."""
    sr = np.sqrt(r)
    rO = r * om
    mag = 1.0 + (pr * pr / (rO * rO)) * (c[0] + c[1] / r + c[2] / (r * sr))
    ph = c[3] * pr / rO + pr ** 3 / rO * c[4]
    return mag * np.cos(ph) + 1j * mag * np.sin(ph)


def imr_template(t_knots, r_knots, phi_knots, pr_knots, mode_series, modes, m1, m2, chi1, chi2, dT, N, amp=1.0,
                 num_add=None, detail=False):
    """One binary's IMR template: sum over modes of the NQC-corrected modes up to
    the attachment sample and of the ringdown after it, on t_j = t_0 + j dT
    (units of M), j < N; returns complex h+ - i... as h+ + i hx (the reference's
    template_channels).  mode_series [2 nm][len]: Re and Im of h_lm e^{i m phi}
    on the knots (as the device's mode knots).  Follows compute_full_waveform
    (eob/seobnrv4phm.py:2308-2556) per binary: the ringdown length is this
    binary's own (the reference pads every binary to the batch's longest), the
    (2,2)-peak time shift is left out (the ringdown depends on t - t_match
    only), and every mode keeps its own real and imaginary part."""
    m1s, m2s = m1 / (m1 + m2), m2 / (m1 + m2)
    L = t_knots.size
    nm = len(modes)
    sph = spline1(t_knots, phi_knots)
    sr = spline1(t_knots, r_knots)
    spr = spline1(t_knots, pr_knots)
    ser = [spline1(t_knots, mode_series[q]) for q in range(2 * nm)]
    t0 = t_knots[0]
    npts = int((t_knots[L - 1] - t0) / dT) + 1
    tg = np.arange(npts) * dT + t0
    tg = np.minimum(tg, t_knots.max())
    om_sparse = seval(sph, t_knots, 1)
    t_peak = t_knots[int(np.argmax(om_sparse))]
    phig = seval(sph, tg, 0)
    omg = seval(sph, tg, 1)
    rg = seval(sr, tg, 0)
    prg = seval(spr, tg, 0)
    hs = []
    for q, (l, m) in enumerate(modes):
        hr = seval(ser[q], tg, 0) + 1j * seval(ser[nm + q], tg, 0)
        hs.append(hr * np.exp(-1j * m * phig))
    hs = np.array(hs)
    dt22 = nr_delta_t22(m1s, m2s, chi1, chi2)
    t_attach = t_peak - dt22
    coefs = []
    for q, (l, m) in enumerate(modes):
        a = np.abs(hs[q])
        ph = np.unwrap(np.angle(hs[q]))
        dTn = dt22 + (10.0 if (l, m) == (5, 5) else 0.0)
        coefs.append(nqc_coefficients(l, m, tg, a, ph, rg, prg, omg, npts, t_peak, dTn, m1s, m2s, chi1, chi2))
    hc = np.array([hs[q] * nqc_factor(coefs[q], rg, prg, omg) for q in range(nm)])
    idx = _argmin_first(np.abs(tg - t_attach))
    dt = tg[1] - tg[0]
    Nw = int(20 / dt) + 1
    t_match = tg[idx]
    mf = final_mass(m1s, m2s, chi1, chi2)
    af = final_spin(m1s, m2s, chi1, chi2)
    w22 = qnm_lal(2, 2, af, mf)
    ringdown_time = int(30 * (1.0 / w22.imag))
    nadd = int(ringdown_time / dt) + 1 if num_add is None else int(num_add)
    eta = _eta(m1s, m2s)
    chi = fit_chi(m1s, m2s, chi1, chi2)
    lo = idx - Nw
    rd = []
    hring = []
    tr = t_match + np.arange(nadd) * dt
    for q, (l, m) in enumerate(modes):
        a = np.abs(hc[q])
        ph = np.unwrap(np.angle(hc[q]))
        tm = t_match - (10.0 if (l, m) == (5, 5) else 0.0)
        sa = spline1(tg[lo:npts], a[lo:npts])
        sp = spline1(tg[lo:npts], ph[lo:npts])
        A0, dA0 = seval(sa, tm, 0)[0], seval(sa, tm, 1)[0]
        phm, wm = seval(sp, tm, 0)[0], seval(sp, tm, 1)[0]
        w = qnm_lal(l, m, af, mf)
        sR, sI = -w.imag, -w.real
        c1f, c2f = rd_fit("c1f", l, m, eta, chi), rd_fit("c2f", l, m, eta, chi)
        if (l, m) == (2, 2) and sR > 2.0 * c1f * math.tanh(c2f):
            c1f = sR / (2.0 * math.tanh(c2f))
        d1f, d2f = rd_fit("d1f", l, m, eta, chi), rd_fit("d2f", l, m, eta, chi)
        c1c = 1 / (c1f * eta) * (dA0 - sR * A0) * math.cosh(c2f) ** 2
        c2c = A0 / eta - 1 / (c1f * eta) * (dA0 - sR * A0) * math.cosh(c2f) * math.sinh(c2f)
        d1c = (wm - sI) * (1 + d2f) / (d1f * d2f)
        x = tr - t_match
        Alm = c1c * np.tanh(c1f * x + c2f) + c2c
        philm = phm - d1c * np.log((1 + d2f * np.exp(-d1f * x)) / (1 + d2f))
        tw = -w.real + 1j * w.imag
        hring.append(eta * Alm * np.exp(1j * philm) * np.exp(1j * tw * x))
        rd.append(dict(amp=A0, damp=dA0, phi=phm, omega=wm, c1c=c1c, c2c=c2c, c1f=c1f, c2f=c2f, d1c=d1c, d1f=d1f, d2f=d2f, qnm=w))
    total = np.zeros(N, dtype=np.complex128)
    upto = min(idx + 1, N)
    total[:upto] = hc[:, :upto].sum(axis=0)
    hr_sum = np.array(hring).sum(axis=0)
    for k in range(1, nadd):
        j = idx + k
        if j < N:
            total[j] = hr_sum[k]
    total *= amp
    if detail:
        return total, dict(npts=npts, t_peak=t_peak, idx=idx, num_add=nadd, nqc=np.array(coefs), rd=rd, mf=mf, af=af,
                           t_match=t_match, modes_raw=hs, modes_nqc=hc, hring=np.array(hring), r=rg, pr=prg, om=omg)
    return total
