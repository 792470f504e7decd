/* Generated by gen_libm_tables.py from mpmath (400 digits). Do not edit. */
#ifndef ST_LIBM_TABLES_H
#define ST_LIBM_TABLES_H

/* pi/2 = P1 + P2 + P3 + P4 + O(2^-210) */
#define ST_PIO2_1 0x1.921fb54442d18p+0 /* 1.5707963267948966 */
#define ST_PIO2_2 0x1.1a62633145c07p-54 /* 6.123233995736766e-17 */
#define ST_PIO2_3 -0x1.f1976b7ed8fbcp-110 /* -1.4973849048591698e-33 */
#define ST_PIO2_4 0x1.4cf98e804177dp-164 /* 5.562271104316826e-50 */
#define ST_TWO_OVER_PI 0x1.45f306dc9c883p-1
#define ST_PIO4 0x1.921fb54442d18p-1
#define ST_INV_LN2 0x1.71547652b82fep+0
#define ST_LN2_HI 0x1.62e42fefa39efp-1
#define ST_LN2_LO 0x1.abc9e3b39803fp-56

#define ST_SIN_1 -0x1.5555555555555p-3 /* (-1)^1/3! */
#define ST_SIN_2 0x1.1111111111111p-7 /* (-1)^2/5! */
#define ST_SIN_3 -0x1.a01a01a01a01ap-13 /* (-1)^3/7! */
#define ST_SIN_4 0x1.71de3a556c734p-19 /* (-1)^4/9! */
#define ST_SIN_5 -0x1.ae64567f544e4p-26 /* (-1)^5/11! */
#define ST_SIN_6 0x1.6124613a86d09p-33 /* (-1)^6/13! */
#define ST_SIN_7 -0x1.ae7f3e733b81fp-41 /* (-1)^7/15! */
#define ST_SIN_8 0x1.952c77030ad4ap-49 /* (-1)^8/17! */
#define ST_COS_1 0x1.5555555555555p-5 /* (-1)^2/4! */
#define ST_COS_2 -0x1.6c16c16c16c17p-10 /* (-1)^3/6! */
#define ST_COS_3 0x1.a01a01a01a01ap-16 /* (-1)^4/8! */
#define ST_COS_4 -0x1.27e4fb7789f5cp-22 /* (-1)^5/10! */
#define ST_COS_5 0x1.1eed8eff8d898p-29 /* (-1)^6/12! */
#define ST_COS_6 -0x1.93974a8c07c9dp-37 /* (-1)^7/14! */
#define ST_COS_7 0x1.ae7f3e733b81fp-45 /* (-1)^8/16! */
#define ST_COS_8 -0x1.6827863b97d97p-53 /* (-1)^9/18! */
#define ST_EXP_2 0x1.0000000000000p-1 /* 1/2! */
#define ST_EXP_3 0x1.5555555555555p-3 /* 1/3! */
#define ST_EXP_4 0x1.5555555555555p-5 /* 1/4! */
#define ST_EXP_5 0x1.1111111111111p-7 /* 1/5! */
#define ST_EXP_6 0x1.6c16c16c16c17p-10 /* 1/6! */
#define ST_EXP_7 0x1.a01a01a01a01ap-13 /* 1/7! */
#define ST_EXP_8 0x1.a01a01a01a01ap-16 /* 1/8! */
#define ST_EXP_9 0x1.71de3a556c734p-19 /* 1/9! */
#define ST_EXP_10 0x1.27e4fb7789f5cp-22 /* 1/10! */
#define ST_EXP_11 0x1.ae64567f544e4p-26 /* 1/11! */
#define ST_EXP_12 0x1.1eed8eff8d898p-29 /* 1/12! */
#define ST_EXP_13 0x1.6124613a86d09p-33 /* 1/13! */
#define ST_EXP_14 0x1.93974a8c07c9dp-37 /* 1/14! */

/* exp: k = rint(x * 128/ln2); ln2/128 = L1 + L2 + L3 + O(2^-150), L1 with 35
   significant bits so that k*L1 is exact for |k| < 2^18 */
#define ST_EXP_INVL 0x1.71547652b82fep+7
#define ST_EXP_L1 0x1.62e42fefc0000p-8
#define ST_EXP_L2 -0x1.c610ca86c3899p-44
#define ST_EXP_L3 0x1.803f2f6af40f3p-99
#define ST_EXPC2_HI 0x1.0000000000000p-1
#define ST_EXPC2_LO 0x0.0p+0
#define ST_EXPC3_HI 0x1.5555555555555p-3
#define ST_EXPC3_LO 0x1.5555555555555p-57
#define ST_EXPC4_HI 0x1.5555555555555p-5
#define ST_EXPC4_LO 0x1.5555555555555p-59
#define ST_EXPC5_HI 0x1.1111111111111p-7
#define ST_EXPC5_LO 0x1.1111111111111p-63
#define ST_EXPC6_HI 0x1.6c16c16c16c17p-10
#define ST_EXPC6_LO -0x1.f49f49f49f49fp-65
/* 2^(j/128) = T[3j] + T[3j+1] + T[3j+2] + O(2^-160), j = 0..127 */
#define ST_EXP_TAB \
  { 0x1.0000000000000p+0, 0x0.0p+0, 0x0.0p+0, \
    0x1.0163da9fb3335p+0, 0x1.b61299ab8cdb7p-54, 0x1.bf48007d80987p-109, \
    0x1.02c9a3e778061p+0, -0x1.19083535b085dp-56, -0x1.9085b0a3d74d5p-110, \
    0x1.04315e86e7f85p+0, -0x1.0a31c1977c96ep-54, -0x1.912fbf44b4040p-112, \
    0x1.059b0d3158574p+0, 0x1.d73e2a475b465p-55, 0x1.05ff94f8d257ep-110, \
    0x1.0706b29ddf6dep+0, -0x1.c91dfe2b13c27p-55, 0x1.fb41f2e2c24abp-110, \
    0x1.0874518759bc8p+0, 0x1.186be4bb284ffp-57, 0x1.15820d96b414fp-111, \
    0x1.09e3ecac6f383p+0, 0x1.1487818316136p-54, -0x1.48b45d1fdc259p-108, \
    0x1.0b5586cf9890fp+0, 0x1.8a62e4adc610bp-54, -0x1.67c9bd6ebf74cp-108, \
    0x1.0cc922b7247f7p+0, 0x1.01edc16e24f71p-54, 0x1.e8aac564e6fe3p-108, \
    0x1.0e3ec32d3d1a2p+0, 0x1.03a1727c57b53p-59, -0x1.5aa76994e9ddbp-113, \
    0x1.0fb66affed31bp+0, -0x1.b9bedc44ebd7bp-57, -0x1.aeb1f49d84259p-112, \
    0x1.11301d0125b51p+0, -0x1.6c51039449b3ap-54, 0x1.9d58b988f562dp-109, \
    0x1.12abdc06c31ccp+0, -0x1.1b514b36ca5c7p-58, -0x1.08d8f42083120p-112, \
    0x1.1429aaea92de0p+0, -0x1.32fbf9af1369ep-54, -0x1.2fe7bb4c76416p-108, \
    0x1.15a98c8a58e51p+0, 0x1.2406ab9eeab0ap-55, -0x1.01b575279c474p-110, \
    0x1.172b83c7d517bp+0, -0x1.19041b9d78a76p-55, 0x1.4f2406aa13ff0p-109, \
    0x1.18af9388c8deap+0, -0x1.11023d1970f6cp-54, 0x1.725f0040b97c5p-110, \
    0x1.1a35beb6fcb75p+0, 0x1.e5b4c7b4968e4p-55, 0x1.ad36183926ae8p-111, \
    0x1.1bbe084045cd4p+0, -0x1.95386352ef607p-54, -0x1.40ca69503718ep-109, \
    0x1.1d4873168b9aap+0, 0x1.e016e00a2643cp-54, 0x1.ea62d0881b918p-110, \
    0x1.1ed5022fcd91dp+0, -0x1.1df98027bb78cp-54, 0x1.e504d36c47475p-108, \
    0x1.2063b88628cd6p+0, 0x1.dc775814a8495p-55, -0x1.781dbc16f1ea4p-111, \
    0x1.21f49917ddc96p+0, 0x1.2a97e9494a5eep-55, -0x1.693c2b3b7106bp-109, \
    0x1.2387a6e756238p+0, 0x1.9b07eb6c70573p-54, -0x1.4d89f9af532e0p-109, \
    0x1.251ce4fb2a63fp+0, 0x1.ac155bef4f4a4p-55, 0x1.1a9c8afdcf797p-112, \
    0x1.26b4565e27cddp+0, 0x1.2bd339940e9d9p-55, 0x1.277393a461b77p-110, \
    0x1.284dfe1f56381p+0, -0x1.a4c3a8c3f0d7ep-54, 0x1.67fdaa2e52d7dp-108, \
    0x1.29e9df51fdee1p+0, 0x1.612e8afad1255p-55, 0x1.de54485604690p-111, \
    0x1.2b87fd0dad990p+0, -0x1.10adcd6381aa4p-59, 0x1.0885fb8796dbdp-113, \
    0x1.2d285a6e4030bp+0, 0x1.0024754db41d5p-54, -0x1.ee9d8f8cb9307p-110, \
    0x1.2ecafa93e2f56p+0, 0x1.1ca0f45d52383p-56, 0x1.d7b08dee6d12ap-111, \
    0x1.306fe0a31b715p+0, 0x1.6f46ad23182e4p-55, 0x1.7b7b2f09cd0d9p-110, \
    0x1.32170fc4cd831p+0, 0x1.a9ce78e18047cp-55, 0x1.b778c882b85e8p-110, \
    0x1.33c08b26416ffp+0, 0x1.32721843659a6p-54, -0x1.406a2ea6cfc6bp-108, \
    0x1.356c55f929ff1p+0, -0x1.b5cee5c4e4628p-55, -0x1.8e524e520d5f2p-109, \
    0x1.371a7373aa9cbp+0, -0x1.63aeabf42eae2p-54, 0x1.87e3e12516bfap-108, \
    0x1.38cae6d05d866p+0, -0x1.e958d3c9904bdp-54, 0x1.0a77a61404f21p-109, \
    0x1.3a7db34e59ff7p+0, -0x1.5e436d661f5e3p-56, 0x1.9b0b1ff17c296p-111, \
    0x1.3c32dc313a8e5p+0, -0x1.efff8375d29c3p-54, -0x1.1143f2a93395ap-109, \
    0x1.3dea64c123422p+0, 0x1.ada0911f09ebcp-55, -0x1.808ba68fa8fb7p-109, \
    0x1.3fa4504ac801cp+0, -0x1.7d023f956f9f3p-54, -0x1.0473e3724200dp-108, \
    0x1.4160a21f72e2ap+0, -0x1.ef3691c309278p-58, -0x1.32b43eafc6518p-114, \
    0x1.431f5d950a897p+0, -0x1.1c7dde35f7999p-55, 0x1.903c496195fefp-109, \
    0x1.44e086061892dp+0, 0x1.89b7a04ef80d0p-59, -0x1.0ac312de3d922p-114, \
    0x1.46a41ed1d0057p+0, 0x1.c944bd1648a76p-54, 0x1.7df404ff21f3ap-108, \
    0x1.486a2b5c13cd0p+0, 0x1.3c1a3b69062f0p-56, 0x1.e1eebae743ac0p-111, \
    0x1.4a32af0d7d3dep+0, 0x1.9cb62f3d1be56p-54, 0x1.91876c761e2c7p-110, \
    0x1.4bfdad5362a27p+0, 0x1.d4397afec42e2p-56, 0x1.c06c7745c2b39p-113, \
    0x1.4dcb299fddd0dp+0, 0x1.8ecdbbc6a7833p-54, 0x1.212c969559b43p-110, \
    0x1.4f9b2769d2ca7p+0, -0x1.4b309d25957e3p-54, -0x1.1aa1fd7b685cdp-112, \
    0x1.516daa2cf6642p+0, -0x1.f768569bd93efp-55, 0x1.90e718226177dp-112, \
    0x1.5342b569d4f82p+0, -0x1.07abe1db13cadp-55, 0x1.fa733951f214cp-111, \
    0x1.551a4ca5d920fp+0, -0x1.d689cefede59bp-55, 0x1.9c991771b0493p-110, \
    0x1.56f4736b527dap+0, 0x1.9bb2c011d93adp-54, -0x1.ff86852a613ffp-111, \
    0x1.58d12d497c7fdp+0, 0x1.295e15b9a1de8p-55, -0x1.a26d92ad1e4c6p-109, \
    0x1.5ab07dd485429p+0, 0x1.6324c054647adp-54, -0x1.744ee506fdafep-109, \
    0x1.5c9268a5946b7p+0, 0x1.c4b1b816986a2p-60, 0x1.ec2735254978cp-119, \
    0x1.5e76f15ad2148p+0, 0x1.ba6f93080e65ep-54, -0x1.95f9ab75fa7d6p-108, \
    0x1.605e1b976dc09p+0, -0x1.3e2429b56de47p-54, -0x1.32c54b92e2588p-110, \
    0x1.6247eb03a5585p+0, -0x1.383c17e40b497p-54, 0x1.5d8e757cfb991p-111, \
    0x1.6434634ccc320p+0, -0x1.c483c759d8933p-55, 0x1.3904000c1c40fp-110, \
    0x1.6623882552225p+0, -0x1.bb60987591c34p-54, 0x1.4a337f4dc0a3bp-108, \
    0x1.68155d44ca973p+0, 0x1.038ae44f73e65p-57, -0x1.f2803633b04ffp-113, \
    0x1.6a09e667f3bcdp+0, -0x1.bdd3413b26456p-54, 0x1.57d3e3adec175p-108, \
    0x1.6c012750bdabfp+0, -0x1.2895667ff0b0dp-56, 0x1.fef5c58766c19p-111, \
    0x1.6dfb23c651a2fp+0, -0x1.bbe3a683c88abp-57, 0x1.a59f88abbe778p-115, \
    0x1.6ff7df9519484p+0, -0x1.83c0f25860ef6p-55, -0x1.001923f4a956ep-110, \
    0x1.71f75e8ec5f74p+0, -0x1.16e4786887a99p-55, -0x1.269796953a4c3p-109, \
    0x1.73f9a48a58174p+0, -0x1.0a8d96c65d53cp-54, 0x1.82ae217f3a768p-108, \
    0x1.75feb564267c9p+0, -0x1.0245957316dd3p-54, -0x1.8f8e7fa19e5e8p-108, \
    0x1.780694fde5d3fp+0, 0x1.866b80a02162dp-54, -0x1.44d42307932f7p-108, \
    0x1.7a11473eb0187p+0, -0x1.41577ee04992fp-55, -0x1.4217a932d10d4p-113, \
    0x1.7c1ed0130c132p+0, 0x1.f124cd1164dd6p-54, -0x1.d4d236cc2bb03p-108, \
    0x1.7e2f336cf4e62p+0, 0x1.05d02ba15797ep-56, 0x1.70a1427f8fcdfp-112, \
    0x1.80427543e1a12p+0, -0x1.27c86626d972bp-54, 0x1.d4e0d71c9b16ep-109, \
    0x1.82589994cce13p+0, -0x1.d4c1dd41532d8p-54, 0x1.0f6ad65cbbac1p-112, \
    0x1.8471a4623c7adp+0, -0x1.8d684a341cdfbp-55, -0x1.591e15c16efd1p-109, \
    0x1.868d99b4492edp+0, -0x1.fc6f89bd4f6bap-54, -0x1.f16f65181d921p-109, \
    0x1.88ac7d98a6699p+0, 0x1.994c2f37cb53ap-54, 0x1.d61283ef385dep-108, \
    0x1.8ace5422aa0dbp+0, 0x1.6e9f156864b27p-54, -0x1.30644a7836333p-110, \
    0x1.8cf3216b5448cp+0, -0x1.0d55e32e9e3aap-56, -0x1.3dab3db839dd6p-111, \
    0x1.8f1ae99157736p+0, 0x1.5cc13a2e3976cp-55, 0x1.3bf26d2b85163p-114, \
    0x1.9145b0b91ffc6p+0, -0x1.dd6792e582524p-54, 0x1.c03855204534ap-109, \
    0x1.93737b0cdc5e5p+0, -0x1.75fc781b57ebcp-57, 0x1.697e257ac0db2p-111, \
    0x1.95a44cbc8520fp+0, -0x1.64b7c96a5f039p-56, -0x1.07053c9a98bbbp-113, \
    0x1.97d829fde4e50p+0, -0x1.d185b7c1b85d1p-54, 0x1.7edb9d7144b6fp-108, \
    0x1.9a0f170ca07bap+0, -0x1.173bd91cee632p-54, -0x1.053987854965fp-110, \
    0x1.9c49182a3f090p+0, 0x1.c7c46b071f2bep-56, 0x1.6376b7943085cp-110, \
    0x1.9e86319e32323p+0, 0x1.824ca78e64c6ep-56, 0x1.0f92c082bbae0p-116, \
    0x1.a0c667b5de565p+0, -0x1.359495d1cd533p-54, 0x1.354084551b4fbp-109, \
    0x1.a309bec4a2d33p+0, 0x1.6305c7ddc36abp-54, 0x1.547fa22c26d17p-108, \
    0x1.a5503b23e255dp+0, -0x1.d2f6edb8d41e1p-54, -0x1.bfd7adfd63f48p-111, \
    0x1.a799e1330b358p+0, 0x1.bcb7ecac563c7p-54, -0x1.678693176f751p-108, \
    0x1.a9e6b5579fdbfp+0, 0x1.0fac90ef7fd31p-54, 0x1.8b16ae39e8cb9p-109, \
    0x1.ac36bbfd3f37ap+0, -0x1.f9234cae76cd0p-55, -0x1.c60dbfc7696f8p-111, \
    0x1.ae89f995ad3adp+0, 0x1.7a1cd345dcc81p-54, 0x1.a7fbc3ae675eap-108, \
    0x1.b0e07298db666p+0, -0x1.bdef54c80e425p-54, 0x1.41cbb95c55600p-109, \
    0x1.b33a2b84f15fbp+0, -0x1.2805e3084d708p-57, 0x1.2babc0edda4d9p-111, \
    0x1.b59728de5593ap+0, -0x1.c71dfbbba6de3p-54, -0x1.c7470081df7dfp-111, \
    0x1.b7f76f2fb5e47p+0, -0x1.5584f7e54ac3bp-56, 0x1.aa64481e1ab72p-111, \
    0x1.ba5b030a1064ap+0, -0x1.efcd30e54292ep-54, -0x1.ad1bf91503c67p-113, \
    0x1.bcc1e904bc1d2p+0, 0x1.23dd07a2d9e84p-55, 0x1.9a164050e1258p-109, \
    0x1.bf2c25bd71e09p+0, -0x1.efdca3f6b9c73p-54, 0x1.27e81cecd59dap-110, \
    0x1.c199bdd85529cp+0, 0x1.11065895048ddp-55, 0x1.99e51125928dap-110, \
    0x1.c40ab5fffd07ap+0, 0x1.b4537e083c60ap-54, 0x1.4a6cdfa70f4f8p-109, \
    0x1.c67f12e57d14bp+0, 0x1.2884dff483cadp-54, -0x1.fc44c329d5cb2p-109, \
    0x1.c8f6d9406e7b5p+0, 0x1.1acbc48805c44p-56, 0x1.6edaac100b8fap-111, \
    0x1.cb720dcef9069p+0, 0x1.503cbd1e949dbp-56, 0x1.d8765566b032ep-110, \
    0x1.cdf0b555dc3fap+0, -0x1.dd83b53829d72p-55, -0x1.aea073a742049p-112, \
    0x1.d072d4a07897cp+0, -0x1.cbc3743797a9cp-54, -0x1.e7044039da0f6p-108, \
    0x1.d2f87080d89f2p+0, -0x1.d487b719d8578p-54, 0x1.2da62b2a9fae7p-111, \
    0x1.d5818dcfba487p+0, 0x1.2ed02d75b3707p-55, -0x1.ab053b05531fcp-111, \
    0x1.d80e316c98398p+0, -0x1.11ec18beddfe8p-54, -0x1.ed04e7ac8765ap-110, \
    0x1.da9e603db3285p+0, 0x1.c2300696db532p-54, 0x1.7f6246f0ec615p-108, \
    0x1.dd321f301b460p+0, 0x1.2da5778f018c3p-54, -0x1.c6cdead661cf3p-108, \
    0x1.dfc97337b9b5fp+0, -0x1.1a5cd4f184b5cp-54, 0x1.b7225a944efd6p-108, \
    0x1.e264614f5a129p+0, -0x1.7b627817a1496p-54, -0x1.b9818808c409ap-108, \
    0x1.e502ee78b3ff6p+0, 0x1.39e8980a9cc8fp-55, 0x1.1e92cb3c2d278p-109, \
    0x1.e7a51fbc74c83p+0, 0x1.2d522ca0c8de2p-54, -0x1.8a757b0b6a9cbp-108, \
    0x1.ea4afa2a490dap+0, -0x1.e9c23179c2893p-54, -0x1.fc0f242bbf3dep-109, \
    0x1.ecf482d8e67f1p+0, -0x1.c93f3b411ad8cp-54, -0x1.0b9dfef44b43bp-108, \
    0x1.efa1bee615a27p+0, 0x1.dc7f486a4b6b0p-54, 0x1.f6dd5d229ff69p-108, \
    0x1.f252b376bba97p+0, 0x1.3a1a5bf0d8e43p-54, 0x1.4c6ad5476b516p-108, \
    0x1.f50765b6e4540p+0, 0x1.9d3e12dd8a18bp-54, -0x1.4019bffc80ef3p-110, \
    0x1.f7bfdad9cbe14p+0, -0x1.dbb12d006350ap-54, 0x1.5c5ce7280fa4dp-108, \
    0x1.fa7c1819e90d8p+0, 0x1.74853f3a5931ep-55, 0x1.dc060c36f7651p-112, \
    0x1.fd3c22b8f71f1p+0, 0x1.2eb74966579e7p-57, 0x1.2f096934ec56cp-111 }

#define ST_LOG_1 0x1.5555555555555p-1 /* 2/3 */
#define ST_LOG_2 0x1.999999999999ap-2 /* 2/5 */
#define ST_LOG_3 0x1.2492492492492p-2 /* 2/7 */
#define ST_LOG_4 0x1.c71c71c71c71cp-3 /* 2/9 */
#define ST_LOG_5 0x1.745d1745d1746p-3 /* 2/11 */
#define ST_LOG_6 0x1.3b13b13b13b14p-3 /* 2/13 */
#define ST_LOG_7 0x1.1111111111111p-3 /* 2/15 */
#define ST_LOG_8 0x1.e1e1e1e1e1e1ep-4 /* 2/17 */
#define ST_LOG_9 0x1.af286bca1af28p-4 /* 2/19 */
#define ST_LOG_10 0x1.8618618618618p-4 /* 2/21 */
#define ST_LOG_11 0x1.642c8590b2164p-4 /* 2/23 */
#define ST_LOG_12 0x1.47ae147ae147bp-4 /* 2/25 */
#define ST_LOG_13 0x1.2f684bda12f68p-4 /* 2/27 */
#define ST_LOG_14 0x1.1a7b9611a7b96p-4 /* 2/29 */
#define ST_SQRT_HALF 0x1.6a09e667f3bcdp-1

/* asin(x) = x + x*t*sum_{n>=1} c_n t^(n-1), t = x^2, c_n = C(2n,n)/(4^n (2n+1)) */
#define ST_ASIN_NCOEF 27
#define ST_ASIN_C0 0x1.5555555555555p-3
#define ST_ASIN_C1 0x1.3333333333333p-4
#define ST_ASIN_C2 0x1.6db6db6db6db7p-5
#define ST_ASIN_C3 0x1.f1c71c71c71c7p-6
#define ST_ASIN_C4 0x1.6e8ba2e8ba2e9p-6
#define ST_ASIN_C5 0x1.1c4ec4ec4ec4fp-6
#define ST_ASIN_C6 0x1.c99999999999ap-7
#define ST_ASIN_C7 0x1.7a87878787878p-7
#define ST_ASIN_C8 0x1.3fde50d79435ep-7
#define ST_ASIN_C9 0x1.12ef3cf3cf3cfp-7
#define ST_ASIN_C10 0x1.df3bd37a6f4dfp-8
#define ST_ASIN_C11 0x1.a6863d70a3d71p-8
#define ST_ASIN_C12 0x1.782dda12f684cp-8
#define ST_ASIN_C13 0x1.51ba308d3dcb1p-8
#define ST_ASIN_C14 0x1.31683bdef7bdfp-8
#define ST_ASIN_C15 0x1.15ee9d45d1746p-8
#define ST_ASIN_C16 0x1.fcaf8fb6db6dbp-9
#define ST_ASIN_C17 0x1.d3d2a8e0dd67dp-9
#define ST_ASIN_C18 0x1.b026f57b13b14p-9
#define ST_ASIN_C19 0x1.90cb77f60c7cep-9
#define ST_ASIN_C20 0x1.750de64d7d05fp-9
#define ST_ASIN_C21 0x1.5c5f56efaaaabp-9
#define ST_ASIN_C22 0x1.464c0950f7d47p-9
#define ST_ASIN_C23 0x1.3275586c5f2f0p-9
#define ST_ASIN_C24 0x1.208d3570ae5a6p-9
#define ST_ASIN_C25 0x1.1052bc5fa960ap-9
#define ST_ASIN_C26 0x1.018f963c229bfp-9

#define ST_TWO_OVER_PI_NWORDS 42
/* word j (j >= 2) holds bits 32(j-2)+1 .. 32(j-2)+32 of the binary fraction 2/pi */
#define ST_TWO_OVER_PI_WORDS \
  { 0x00000000u, 0x00000000u, 0xa2f9836eu, 0x4e441529u, 0xfc2757d1u, 0xf534ddc0u, \
    0xdb629599u, 0x3c439041u, 0xfe5163abu, 0xdebbc561u, 0xb7246e3au, 0x424dd2e0u, \
    0x06492eeau, 0x09d1921cu, 0xfe1deb1cu, 0xb129a73eu, 0xe88235f5u, 0x2ebb4484u, \
    0xe99c7026u, 0xb45f7e41u, 0x3991d639u, 0x835339f4u, 0x9c845f8bu, 0xbdf9283bu, \
    0x1ff897ffu, 0xde05980fu, 0xef2f118bu, 0x5a0a6d1fu, 0x6d367ecfu, 0x27cb09b7u, \
    0x4f463f66u, 0x9e5fea2du, 0x7527bac7u, 0xebe5f17bu, 0x3d0739f7u, 0x8a5292eau, \
    0x6bfb5fb1u, 0x1f8d5d08u, 0x56033046u, 0xfc7b6babu, 0xf0cfbc20u, 0x9af4361du }

#endif
