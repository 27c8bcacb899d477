"""Plain AES-128 block encryption (FIPS-197), used only to recover the GHASH
subkey and tag mask from the NIST GCM known-answer vectors. It works in
GF(2^8) under a different polynomial and shares no code with what is tested."""


def _xtime(b):
    b <<= 1
    return (b ^ 0x11B) & 0xFF if b & 0x100 else b


def _gmul(a, b):
    r = 0
    while b:
        if b & 1:
            r ^= a
        a = _xtime(a)
        b >>= 1
    return r


def _sbox():
    box = [0] * 256
    for x in range(256):
        inv = 0
        if x:
            for y in range(1, 256):
                if _gmul(x, y) == 1:
                    inv = y
                    break
        s = inv
        for k in range(1, 5):
            s ^= ((inv << k) | (inv >> (8 - k))) & 0xFF
        box[x] = s ^ 0x63
    return box


SBOX = _sbox()


def _expand(key):
    w = [list(key[4 * i:4 * i + 4]) for i in range(4)]
    rcon = 1
    for i in range(4, 44):
        t = list(w[i - 1])
        if i % 4 == 0:
            t = t[1:] + t[:1]
            t = [SBOX[b] for b in t]
            t[0] ^= rcon
            rcon = _xtime(rcon)
        w.append([a ^ b for a, b in zip(w[i - 4], t)])
    return [sum(w[4 * r:4 * r + 4], []) for r in range(11)]


def encrypt_block(key: bytes, block: bytes) -> bytes:
    rk = _expand(key)
    s = [b ^ k for b, k in zip(block, rk[0])]
    for rnd in range(1, 11):
        s = [SBOX[b] for b in s]
        # ShiftRows on column-major state: byte (r, c) at index 4c + r
        s = [s[(4 * ((c + r) % 4)) + r] for c in range(4) for r in range(4)]
        s = [s[4 * (i // 4) + (i % 4)] for i in range(16)]
        if rnd != 10:
            out = []
            for c in range(4):
                a = s[4 * c:4 * c + 4]
                out += [
                    _gmul(a[0], 2) ^ _gmul(a[1], 3) ^ a[2] ^ a[3],
                    a[0] ^ _gmul(a[1], 2) ^ _gmul(a[2], 3) ^ a[3],
                    a[0] ^ a[1] ^ _gmul(a[2], 2) ^ _gmul(a[3], 3),
                    _gmul(a[0], 3) ^ a[1] ^ a[2] ^ _gmul(a[3], 2),
                ]
            s = out
        s = [b ^ k for b, k in zip(s, rk[rnd])]
    return bytes(s)
