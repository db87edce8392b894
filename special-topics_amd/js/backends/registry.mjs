// Backend registry: lazy loaders, availability probing and selection, with the
// behaviour of js_fdtd/src/backends/registry.js:13-84 -- a backend whose module
// fails to load, or whose isAvailable() throws, is reported unavailable and skipped;
// an explicit preference is tried first, then the default order.
//
// This package ships one backend, 'hip'.  Hosts that also carry the reference's
// engines add them with registerBackend(name, loader) (e.g. 'wasm-cpu').

const LOADERS = {
  hip: () => import('./hipBackend.mjs').then((m) => m.default),
};

const DEFAULT_ORDER = ['hip'];
const cache = new Map();

/** Add (or replace) a backend loader: loader() resolves to a backend object. */
export function registerBackend(name, loader, { priority = DEFAULT_ORDER.length } = {}) {
  LOADERS[name] = loader;
  cache.delete(name);
  if (!DEFAULT_ORDER.includes(name)) DEFAULT_ORDER.splice(priority, 0, name);
}

export function knownBackends() {
  return Object.keys(LOADERS);
}

/** The backend object, or null when its module fails to load. */
export async function loadBackend(name) {
  if (cache.has(name)) return cache.get(name);
  const loader = LOADERS[name];
  if (!loader) throw new Error(`Unknown backend "${name}". Known: ${Object.keys(LOADERS).join(', ')}`);
  let backend = null;
  try {
    backend = await loader();
  } catch (err) {
    console.warn(`[backends] backend "${name}" failed to load: ${err.message}`);
  }
  cache.set(name, backend);
  return backend;
}

/** { name, backend } of the first available backend: `preferred` first unless 'auto'. */
export async function selectBackend(preferred = 'auto') {
  const order = preferred && preferred !== 'auto' ? [preferred, ...DEFAULT_ORDER] : DEFAULT_ORDER.slice();
  const seen = new Set();
  for (const name of order) {
    if (seen.has(name)) continue;
    seen.add(name);
    const backend = await loadBackend(name);
    if (!backend) continue;
    let ok = false;
    try {
      ok = await backend.isAvailable();
    } catch (err) {
      ok = false;
    }
    if (ok) return { name, backend };
    if (preferred === name) console.warn(`[backends] requested "${name}" is unavailable; falling back.`);
  }
  throw new Error(`No FDTD backend is available (tried ${[...seen].join(', ')}).`);
}

/** [{name, available, detail?}] for every registered backend. */
export async function listBackends() {
  const out = [];
  for (const name of Object.keys(LOADERS)) {
    const backend = await loadBackend(name);
    if (!backend) {
      out.push({ name, available: false, detail: 'module failed to load' });
      continue;
    }
    let available = false;
    let detail;
    try {
      available = await backend.isAvailable();
    } catch (err) {
      detail = err.message;
    }
    out.push({ name, available, detail });
  }
  return out;
}
