"""AV1's symbol decoder restated from the specification (8.2.2 init_symbol, 8.2.6
read_symbol with the renormalization, 8.2.4 exit_symbol's trailing-bit check), without CDF
adaptation (disable_cdf_update = 1).  Test infrastructure: decodes what the library's range
coder writes."""


class SymbolDecoder:
    def __init__(self, data: bytes):
        self.data = data
        self.pos = 0  # bit position
        sz = len(data)
        num_bits = min(sz * 8, 15)
        buf = self.f(num_bits)
        padded = buf << (15 - num_bits)
        self.value = ((1 << 15) - 1) ^ padded
        self.rng = 1 << 15
        self.max_bits = 8 * sz - 15

    def f(self, n):
        v = 0
        for _ in range(n):
            byte = self.data[self.pos >> 3] if (self.pos >> 3) < len(self.data) else 0
            v = (v << 1) | ((byte >> (7 - (self.pos & 7))) & 1)
            self.pos += 1
        return v

    def symbol(self, cdf, n):
        """cdf: the n - 1 cumulative values (32768 * P(symbol <= k))."""
        full = [int(c) for c in cdf] + [32768]
        cur = self.rng
        s = -1
        while True:
            s += 1
            prev = cur
            f = (1 << 15) - full[s]
            cur = ((self.rng >> 8) * (f >> 6) >> 1) + 4 * (n - s - 1)
            if self.value >= cur:
                break
        self.rng = prev - cur
        self.value -= cur
        bits = 15 - (self.rng.bit_length() - 1)
        self.rng <<= bits
        num_bits = min(bits, max(0, self.max_bits))
        new = self.f(num_bits)
        padded = new << (bits - num_bits)
        self.value = padded ^ (((self.value + 1) << bits) - 1)
        self.max_bits -= bits
        return s

    def boolean(self, cdf0):
        return self.symbol([cdf0], 2)

    def trailing_ok(self):
        """exit_symbol: the bit at trailingBitPosition is 1 and every later bit of the tile 0."""
        if self.max_bits < -14:
            return False
        trailing = self.pos - min(15, self.max_bits + 15)
        total = len(self.data) * 8
        bits = [((self.data[p >> 3] >> (7 - (p & 7))) & 1) for p in range(trailing, total)]
        return len(bits) > 0 and bits[0] == 1 and not any(bits[1:])
