"""Minimal INI configuration with the accessor API of frfs/inifile.py:19-141
(``get``, ``getint``, ``getfloat`` with defaults, ``set``, ``hasopt``)."""
import configparser


class Inifile:
    def __init__(self, inistr=None):
        self._cp = configparser.ConfigParser(inline_comment_prefixes=(';',))
        self._cp.optionxform = str
        if inistr:
            self._cp.read_string(inistr)

    @classmethod
    def load(cls, path):
        with open(path) as fh:
            return cls(fh.read())

    @classmethod
    def from_dict(cls, d):
        ini = cls()
        for sec, kv in d.items():
            for k, v in kv.items():
                ini.set(sec, k, v)
        return ini

    def set(self, section, option, value):
        if not self._cp.has_section(section):
            self._cp.add_section(section)
        self._cp.set(section, option, str(value))

    def hasopt(self, section, option):
        return self._cp.has_option(section, option)

    _NO_DEFAULT = object()

    def get(self, section, option, default=_NO_DEFAULT):
        if self._cp.has_option(section, option):
            return self._cp.get(section, option)
        if default is self._NO_DEFAULT:
            raise KeyError(f'[{section}] {option}')
        self.set(section, option, default)
        return str(default)

    def getint(self, section, option, default=_NO_DEFAULT):
        return int(self.get(section, option, default))

    def getfloat(self, section, option, default=_NO_DEFAULT):
        return float(self.get(section, option, default))

    def tostr(self):
        import io
        buf = io.StringIO()
        self._cp.write(buf)
        return buf.getvalue()
